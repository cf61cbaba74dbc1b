// tdibc_kernels.h -- time-domain impedance boundary controller on the device.
//
// Restates cases/nit/nit.c:13-47 (explicit-Euler ADE of a multipole admittance,
// updated in place so the chi update sees the new psi, as nit.c:31 calls it with
// aliased in/out arrays) and the per-step driver glue of nit.c:273-303 (wall
// pressure from the lagged density, velocity scaled into lattice units, applied
// to the active rows of the driven boundary).  FP64 throughout; the lumped mode
// sums the probe line in index order, as nit.c:277-280 does.
#pragma once
#include "step_kernels.h"

namespace lbmx {

struct TdibcParams {
    int which, probe_is_column, probe_index, per_node, record_node, capacity, nedge;
    double rho0, cs, cp, dt, cu, gain;
    double Zinf, A[2], lambdas[2], B[2], C[2], alpha[2], beta[2];
};

struct TdibcState {
    double phi[2], psi[2], chi[2];
};

struct TdibcArgs {
    Geo g;
    StepFlags s;
    Fields F;
    TdibcParams p;
    TdibcState *st;      // one (lumped) or nedge (per node)
    double *last;        // (p', v) of the previous step, as nit.c:273 prints them
    const int *active;   // nedge flags
    double *val;         // driven boundary value array (nedge)
    double *hist;        // capacity x 3
    long long step;      // history row of this step
};

__device__ __forceinline__ double ade_velocity(TdibcState &s, const TdibcParams &P, double p)
{
    for (int i = 0; i < 2; i++) {
        s.phi[i] = s.phi[i] + P.dt * (p + P.lambdas[i] * s.phi[i]);
        s.psi[i] = s.psi[i] + P.dt * (p + P.alpha[i] * s.psi[i] - P.beta[i] * s.chi[i]);
        s.chi[i] = s.chi[i] + P.dt * (P.beta[i] * s.psi[i] + P.alpha[i] * s.chi[i]);
    }
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < 2; i++) {
        a += P.A[i] * s.phi[i];
        b += P.B[i] * s.psi[i];
        c += P.C[i] * s.chi[i];
    }
    return P.Zinf * p - a - 2.0 * (b + c);
}

__device__ __forceinline__ size_t probe_node(const TdibcArgs &a, int e)
{
    return a.p.probe_is_column ? (size_t)e * a.g.pitch + a.p.probe_index
                               : (size_t)a.p.probe_index * a.g.pitch + e;
}

constexpr int kTdibcThreads = 1024;

template <typename T>
__global__ void __launch_bounds__(kTdibcThreads) k_tdibc_lumped(const T *__restrict__ S,
                                                                TdibcArgs a)
{
    __shared__ double buf[kTdibcThreads];
    __shared__ double value;
    const int n = a.p.nedge;
    double acc = 0.0;
    for (int base = 0; base < n; base += kTdibcThreads) {
        const int e = base + threadIdx.x;
        if (e < n) buf[threadIdx.x] = lagged_at(S, a.g, a.F, a.s, probe_node(a, e)).rho;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = (n - base < kTdibcThreads) ? n - base : kTdibcThreads;
            for (int t = 0; t < m; t++) acc += buf[t];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const TdibcParams &P = a.p;
        double rs = acc;
        rs /= n;
        const double p = P.cs * P.cs * (rs - P.rho0) * P.cp;
        if (a.step < P.capacity) {
            a.hist[3 * a.step + 0] = a.step * P.dt;
            a.hist[3 * a.step + 1] = a.last[0];
            a.hist[3 * a.step + 2] = a.last[1];
        }
        TdibcState st = a.st[0];
        const double v = ade_velocity(st, P, p);
        a.st[0] = st;
        a.last[0] = p;
        a.last[1] = v;
        value = P.gain * (v / P.cu);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x)
        a.val[e] = (!a.active || a.active[e]) ? value : 0.0;
}

template <typename T>
__global__ void __launch_bounds__(256) k_tdibc_nodes(const T *__restrict__ S, TdibcArgs a)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.p.nedge) return;
    const TdibcParams &P = a.p;
    const bool on = !a.active || a.active[e];
    if (!on) {
        a.val[e] = 0.0;
        return;
    }
    const double rho = lagged_at(S, a.g, a.F, a.s, probe_node(a, e)).rho;
    const double p = P.cs * P.cs * (rho - P.rho0) * P.cp;
    TdibcState st = a.st[e];
    const double v = ade_velocity(st, P, p);
    a.st[e] = st;
    a.val[e] = P.gain * (v / P.cu);
    if (e == P.record_node) {
        if (a.step < P.capacity) {
            a.hist[3 * a.step + 0] = a.step * P.dt;
            a.hist[3 * a.step + 1] = a.last[0];
            a.hist[3 * a.step + 2] = a.last[1];
        }
        a.last[0] = p;
        a.last[1] = v;
    }
}

// Density probes: lagged density at listed nodes, one history row per step.
struct ProbeArgs {
    Geo g;
    StepFlags s;
    Fields F;
    const long long *nodes;
    int n, capacity;
    double *hist;
    long long step;
};

template <typename T>
__global__ void __launch_bounds__(256) k_probes(const T *__restrict__ S, ProbeArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n || a.step >= a.capacity) return;
    a.hist[a.step * a.n + i] = lagged_at(S, a.g, a.F, a.s, (size_t)a.nodes[i]).rho;
}

}  // namespace lbmx
