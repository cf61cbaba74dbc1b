/*
 * lbm.h -- umbrella header of the lbc C API (reference include/lbm.h:1-9).
 * Link the reference's case drivers against lbc_amd/_lib/liblbm_hip.so instead of
 * compiling src/ (gcc driver.c -I include -L lbc_amd/_lib -llbm_hip -lm).
 */
#ifndef LBM_H
#define LBM_H

#include "lbm_types.h"
#include "lbm_core.h"
#include "lbm_bc.h"
#include "lbm_utils.h"

#endif
