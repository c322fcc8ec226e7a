// Triton shallow-water flood model: the step chain of Triton/src/triton_appfis.hpp:
// 316-386 with the functors of triton_appfis.hpp:567-1112, plus the config's uniform
// rainfall (added as h += rain*dt right after the update, where the reference adds its
// point sources), fused into two kernels per step:
//
//   K1 (flux + update + rain): a wave marches up a strip of 62 columns x 32 rows, one
//       lane per column, solving each edge once (X edges with the left cell's values
//       shifted in across lanes, Y edges from the row below kept in registers); every
//       cell accumulates rhs*0 / rhs*1 from its four edges in the order of the
//       reference's X-then-Y functors, applies the update with implicit Manning
//       friction and the rain, and writes the new state to the other buffer of a
//       ping-pong pair (the reference's reset pass and six rhs arrays never touch
//       memory).
//   K2 (wet/dry fix + CFL + volume): the momentum fix of triton_appfis.hpp:1095-1112 in
//       place, then the CFL candidate of triton_appfis.hpp:567-586 from the fixed state
//       (wave min) and a compensated FP64 sum of h, written as one partial per strip --
//       no atomics (per-scenario atomics serialise hundreds of thousands of waves on a
//       handful of addresses).
//   K3 (finalise): one block per scenario combines its strip partials in a fixed order
//       and does on the device what the reference does on the host
//       (triton_appfis.hpp:322-335, 381-385): simtime += dt, the print_interval
//       bookkeeping, the next dt (fallback when every cell is dry, clamp to the next
//       print boundary), and records time / volume / dt of the step.
//
// No host round trip per step: the reference's Execute2DReduce(MIN) + D2H + 3 host
// syncs per step become device-side state, so a chunk of steps can be captured in a
// hipGraph and replayed.  Ensembles: every phase is one launch over a work-list of
// strips spanning all scenarios (the scenario is found by binary search over
// per-scenario offsets), with per-scenario dt and reductions.
//
// REAL = double (reference dtype) or float (FP32 bulk form; the volume is still a
// compensated FP64 sum).  This header holds the kernels and the host driver run_flood<R>
// in an anonymous namespace; two translation units instantiate it: flood.hip (the C
// entry points and the FP64 form, compiled without FMA contraction, so that the
// operation order is the functors') and flood_f32.hip (the FP32 bulk form, compiled with
// multiply-adds contracted into FMAs: its accuracy is a stated tolerance against FP64,
// and the contraction is 3% of its step time).
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cfloat>
#include <vector>

using namespace appfis;

namespace {

constexpr int BT = 256;  // threads per block: four waves, one strip each
constexpr int CHUNK = 20;  // steps per captured graph (even: buffer parity returns)

template <typename R> struct RealLim;
template <> struct RealLim<double> {
    __host__ __device__ static double max() { return DBL_MAX; }
};
template <> struct RealLim<float> {
    __host__ __device__ static float max() { return FLT_MAX; }
};

struct ScenState {
    double dt;       // dt of the step about to run
    double simtime;  // simulated time (s)
    int print_id;
    int step;        // steps completed in this call
    int src_low;     // hydrograph row where the next interval search starts
};

template <typename R>
struct Scen {
    R* h[2];
    R* qx[2];
    R* qy[2];
    const R* z;
    const R* n2;
    int nx, ny;
    int strips_x;      // strips of MW x MR cells (every kernel of the step works on them)
    int strip0;        // first global strip of this scenario
    int nstrips;
    ScenState* st;
    double* partials;  // 3 doubles per strip: CFL minimum, compensated volume sum and its error
    double* rec_time;  // per-step records (nsteps each) or null
    double* rec_vol;
    double* rec_dt;
    // point sources: nsrc padded cell indices, a hydrograph of src_rows times (s) and
    // src_rows x nsrc flows (m^3/s, row-major)
    int nsrc, src_rows;
    const int* src_cell;
    const double* src_time;
    const double* src_flow;
};

struct Params {
    double dx, cn, hextra, dt_fallback, print_interval, rain;
    int nsteps_rec;  // length of the record arrays
    int fixed_dt;    // time_increment_fixed: every step takes dt_fallback
    double end_time; // > 0: a step starting at simtime >= end_time does nothing
};

// The reference loops while simtime < sim_duration (triton_appfis.hpp:316): with an end
// time, every kernel of a step that would start at or past it returns at once (the test
// reads a value the previous launch wrote, the same for every thread of the launch).
__device__ __forceinline__ bool finished(const ScenState* st, const Params& P)
{
    return P.end_time > 0.0 && st->simtime >= P.end_time;
}

// Division and square root: correctly rounded for R = double (the reference's
// arithmetic: the FP64 form reproduces the oracle); for the FP32 bulk form the hardware
// reciprocal and square-root approximations (v_rcp_f32 / v_sqrt_f32, about 1 ulp), well
// inside that form's stated error against FP64.
template <typename R> __device__ __forceinline__ R fdiv(R a, R b) { return a / b; }
template <> __device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R> __device__ __forceinline__ R fsqrt(R a) { return sqrt(a); }
template <> __device__ __forceinline__ float fsqrt<float>(float a) { return __builtin_amdgcn_sqrtf(a); }
// Cube root of a positive depth: libm's for R = double; 2^(log2(a)/3) from the hardware
// log2 / exp2 approximations for the FP32 form (a few ulp).
template <typename R> __device__ __forceinline__ R fcbrt(R a) { return cbrt(a); }
template <> __device__ __forceinline__ float fcbrt<float>(float a)
{
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(a) * (1.0f / 3.0f));
}

// dt * d / dx of an edge's contribution: the functors' expression for R = double; one
// multiply by a per-thread dt/dx for the FP32 form.
template <typename R>
struct RhsScale {
    R dt, dx;
    __device__ RhsScale(R dt_, R dx_) : dt(dt_), dx(dx_) {}
    __device__ R operator()(R d) const { return fdiv(dt * d, dx); }
};
template <>
struct RhsScale<float> {
    float f;
    __device__ RhsScale(float dt_, float dx_) : f(dt_ * __builtin_amdgcn_rcpf(dx_)) {}
    __device__ float operator()(float d) const { return d * f; }
};

// Edge normal: (1, 0) for DIR 0 (X edges), (0, 1) for DIR 1 (Y edges).  The FP64 form
// multiplies the normal out exactly as the reference functors do (a product with the
// zero component is +-0 and cannot be dropped without changing signed zeros, so the
// results stay bit-identical); the FP32 bulk form drops the terms multiplied by the
// zero component.
template <typename R, int DIR>
struct Normal {
    static constexpr bool EXACT = sizeof(R) == 8;
    static constexpr R nx = DIR == 0 ? (R)1 : (R)0, ny = DIR == 0 ? (R)0 : (R)1;
    __device__ static R dot(R a, R b)  // a*nx + b*ny
    {
        if constexpr (EXACT) return a * nx + b * ny;
        else return DIR == 0 ? a : b;
    }
    __device__ static R mx(R c)  // c*nx
    {
        if constexpr (EXACT) return c * nx;
        else return DIR == 0 ? c : (R)0;
    }
    __device__ static R my(R c)  // c*ny
    {
        if constexpr (EXACT) return c * ny;
        else return DIR == 1 ? c : (R)0;
    }
    __device__ static R addx(R a, R c)  // a + c*nx
    {
        if constexpr (EXACT) return a + c * nx;
        else return DIR == 0 ? a + c : a;
    }
    __device__ static R addy(R a, R c)  // a + c*ny
    {
        if constexpr (EXACT) return a + c * ny;
        else return DIR == 1 ? a + c : a;
    }
    __device__ static R subx(R a, R c)  // a - c*nx
    {
        if constexpr (EXACT) return a - c * nx;
        else return DIR == 0 ? a - c : a;
    }
    __device__ static R suby(R a, R c)  // a - c*ny
    {
        if constexpr (EXACT) return a - c * ny;
        else return DIR == 1 ? a - c : a;
    }
};

// Roe flux with entropy fix, bed-slope source and wet/dry treatment -- the shared body
// of the X functor (triton_appfis.hpp:608-824, DIR 0) and the Y functor (827-1043, DIR 1).
// Cell velocity of the functors (triton_appfis.hpp:620-631): q/h where h >= hextra, else 0.
template <typename R> __device__ inline void velocity(R h, R qx, R qy, R hextra, R& u, R& v)
{
    if (h >= hextra) {
        u = fdiv(qx, h);
        v = fdiv(qy, h);
    } else {
        u = (R)0.0;
        v = (R)0.0;
    }
}

// The edge solve takes each side's velocities (u, v) and sqrt(h) precomputed per cell:
// the values are the functors' own, computed once per cell instead of once per edge.
template <typename R, int DIR>
__device__ inline bool roe(R h1, R h2, R qx1, R qy1, R z1, R qx2, R qy2, R z2, R sqh1, R sqh2, R u1, R v1, R u2, R v2,
                           R dx, R dt, R d1[3], R d2[3])
{
    using N = Normal<R, DIR>;
    const R G = (R)9.81, SQRTG = (R)3.132091953, EPS12 = (R)1e-12;
    if (!(h1 > (R)0.0 || h2 > (R)0.0)) return false;
    R alpha[3], beta[3], eigen[3], eigenE[3], eigev[3][3];
    const R h12 = (R)0.5 * (h1 + h2);
    const R dh = h2 - h1;
    const R c12 = fsqrt(G * h12);
    const R u12 = fdiv(u1 * sqh1 + u2 * sqh2, sqh1 + sqh2);
    const R v12 = fdiv(v1 * sqh1 + v2 * sqh2, sqh1 + sqh2);
    const R un = N::dot(u12, v12);
    eigen[0] = un - c12;
    eigen[1] = un;
    eigen[2] = un + c12;
#pragma unroll
    for (int i = 0; i < 3; i++) eigenE[i] = (R)0.0;
    R e1 = N::dot(u1, v1) - SQRTG * sqh1;
    R e2 = N::dot(u2, v2) - SQRTG * sqh2;
    if (e1 < (R)0.0 && e2 > (R)0.0) {
        eigenE[0] = eigen[0] - fdiv(e1 * (e2 - eigen[0]), e2 - e1);
        eigen[0] = fdiv(e1 * (e2 - eigen[0]), e2 - e1);
    }
    e1 = N::dot(u1, v1) + SQRTG * sqh1;
    e2 = N::dot(u2, v2) + SQRTG * sqh2;
    if (e1 < (R)0.0 && e2 > (R)0.0) {
        eigenE[2] = eigen[2] - fdiv(e2 * (eigen[2] - e1), e2 - e1);
        eigen[2] = fdiv(e2 * (eigen[2] - e1), e2 - e1);
    }
    eigev[0][0] = (R)1.0;
    eigev[0][1] = N::subx(u12, c12);
    eigev[0][2] = N::suby(v12, c12);
    eigev[1][0] = (R)0.0;
    eigev[1][1] = N::my(-c12);
    eigev[1][2] = N::mx(c12);
    eigev[2][0] = (R)1.0;
    eigev[2][1] = N::addx(u12, c12);
    eigev[2][2] = N::addy(v12, c12);
    alpha[0] = (R)0.5 * (dh - fdiv(N::dot(qx2 - qx1, qy2 - qy1) - un * dh, c12));
    alpha[1] = fdiv(N::mx((qy2 - qy1) - v12 * dh) - N::my((qx2 - qx1) - u12 * dh), c12);
    alpha[2] = (R)0.5 * (dh + fdiv(N::dot(qx2 - qx1, qy2 - qy1) - un * dh, c12));
    const R dz = z2 - z1;
    R l1 = z1 + h1;
    R l2 = z2 + h2;
    R dzp = dz;
    R hp;
    if (dz >= (R)0.0) {
        hp = h1;
        if (l1 < z2) dzp = h1;
    } else {
        hp = h2;
        if (l2 < z1) dzp = -h2;
    }
    beta[0] = fdiv((R)0.5 * G * (hp - (R)0.5 * fabs(dzp)) * dzp, c12);
    hp = h1 + alpha[0];
    if (eigen[0] * eigen[2] < (R)0.0 && hp > (R)0.0 && h1 > (R)0.0 && h2 > (R)0.0) {
        beta[0] = fmax(beta[0], alpha[0] * eigen[0] - fdiv(h1 * dx * (R)0.5, dt));
        beta[0] = fmin(beta[0], -alpha[2] * eigen[2] + fdiv(h2 * dx * (R)0.5, dt));
    }
    beta[1] = (R)0.0;
    beta[2] = -beta[0];
#pragma unroll
    for (int i = 0; i < 3; i++) { d1[i] = (R)0.0; d2[i] = (R)0.0; }
    l1 = hp - fdiv(beta[0], eigen[0]);
    l2 = hp + fdiv(beta[2], eigen[2]);
    if (l2 < -EPS12 && h2 < EPS12) {
        d1[0] += (eigen[0] * alpha[0] - beta[0]) + (eigen[2] * alpha[2] - beta[2]);
    } else if (l1 < -EPS12 && h1 < EPS12) {
        d2[0] += (eigen[0] * alpha[0] - beta[0]) + (eigen[2] * alpha[2] - beta[2]);
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (eigen[i] > (R)0.0) {
                d2[0] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][0];
                d2[1] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][1];
                d2[2] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][2];
            } else {
                d1[0] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][0];
                d1[1] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][1];
                d1[2] += (eigen[i] * alpha[i] - beta[i]) * eigev[i][2];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (eigenE[i] > (R)0.0) {
            d2[0] += (eigenE[i] * alpha[i]) * eigev[i][0];
            d2[1] += (eigenE[i] * alpha[i]) * eigev[i][1];
            d2[2] += (eigenE[i] * alpha[i]) * eigev[i][2];
        } else {
            d1[0] += (eigenE[i] * alpha[i]) * eigev[i][0];
            d1[1] += (eigenE[i] * alpha[i]) * eigev[i][1];
            d1[2] += (eigenE[i] * alpha[i]) * eigev[i][2];
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (fabs(d1[i]) < EPS12) d1[i] = (R)0.0;
        if (fabs(d2[i]) < EPS12) d2[i] = (R)0.0;
    }
    return true;
}

template <typename R> __device__ inline R sqrt_pos(R h) { return h > (R)0.0 ? fsqrt(h) : (R)0.0; }

// scenario of a global strip: binary search over strip0 (uniform per wave)
template <typename R>
__device__ inline int find_scen(const Scen<R>* sc, int nscen, int strip)
{
    int lo = 0, hi = nscen - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sc[mid].strip0 <= strip) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One cell's state and the per-cell values every edge solve of it uses.
template <typename R>
struct CellV {
    R h, qx, qy, z, n2, sq, u, v;
};

// Unconditional loads (the compiler can then count them in flight): callers clamp the
// offset into the array, and the values of lanes or rows outside it are never used.
template <typename R>
__device__ inline void cell_load(CellV<R>& c, const gmem<R>* h, const gmem<R>* qx, const gmem<R>* qy, const gmem<R>* z,
                                 const gmem<R>* n2, size_t off)
{
    c.h = h[off];
    c.qx = qx[off];
    c.qy = qy[off];
    c.z = z[off];
    c.n2 = n2[off];
}

template <typename R>
__device__ inline void cell_derive(CellV<R>& c, R hextra)
{
    c.sq = sqrt_pos(c.h);
    velocity(c.h, c.qx, c.qy, hextra, c.u, c.v);
}

// K1 as a row march.  A wave owns a strip of MW columns x MR rows; lane l holds column
// i0-1+l, so lanes 1..MW are the strip's cells and lanes 0 and 63 its halo columns.  The
// wave walks the rows j0-1 .. j0+MR keeping two rows in registers:
//   X edge of row j between columns (c-1, c): solved by the lane of c, with the left
//     cell's values shifted in from lane l-1; the cell on the right of the edge takes
//     rhs*1, the left cell's rhs*0 is shifted back to lane l-1;
//   Y edge between rows j-1 (side 2) and j (side 1): solved by each cell lane from its
//     two rows; row j takes rhs*0 +=, row j-1 rhs*1 +=, after which row j-1 is complete
//     and is updated (implicit Manning friction, rain) into the other buffer.
// Every edge is solved once; each cell's velocities and sqrt(h) once; the arguments of
// every solve and the order of the rhs accumulation are those of the functors, so the
// results are those of the per-cell form.  No LDS and no barriers.
constexpr int MW = 62;            // strip columns (lanes 1..62)
constexpr int MR = 32;            // strip rows
constexpr int MWAVES = BT / 64;   // strips per block

template <typename R>
__global__ __launch_bounds__(BT) void k_flux_march(const Scen<R>* __restrict__ sc, int nscen, int nstrips, int par,
                                                   Params P)
{
    const int lane = threadIdx.x & 63;
    const int strip = blockIdx.x * MWAVES + (threadIdx.x >> 6);  // uniform per wave
    if (strip >= nstrips) return;
    const int s = find_scen(sc, nscen, strip);
    const Scen<R>& S = sc[s];
    if (finished(S.st, P)) return;
    const int t = strip - S.strip0;
    const int ty = t / S.strips_x, tx = t - ty * S.strips_x;
    const int nx = S.nx, ny = S.ny;
    const size_t W = (size_t)nx + 2;
    const int col = tx * MW + lane;  // padded column of this lane
    const int j0 = ty * MR + 1;
    const bool colmem = col <= nx + 1;
    const bool cell = lane >= 1 && lane <= MW && col <= nx;
    const bool xedge = lane >= 1 && colmem;  // edge (col-1, col) with col-1 <= nx
    const gmem<R>* h = to_global((const R*)S.h[par]);
    const gmem<R>* qx = to_global((const R*)S.qx[par]);
    const gmem<R>* qy = to_global((const R*)S.qy[par]);
    const gmem<R>* z = to_global(S.z);
    const gmem<R>* n2 = to_global(S.n2);
    const int o = par ^ 1;  // the buffer written
    gmem<R>* hout = to_global(S.h[o]);
    gmem<R>* qxout = to_global(S.qx[o]);
    gmem<R>* qyout = to_global(S.qy[o]);
    const R dt = (R)S.st->dt;
    const R dx = (R)P.dx, hextra = (R)P.hextra, rain = (R)P.rain;
    const R G = (R)9.81, EPS12 = (R)1e-12;
    const int rend = min(MR, ny + 1 - j0);  // last row offset: rows up to ny+1 exist
    const RhsScale<R> scale(dt, dx);

    CellV<R> prev, cur, nxt;
    const size_t colc = (size_t)(colmem ? col : nx + 1);  // lanes past the array load column nx+1
    cell_load(prev, h, qx, qy, z, n2, colc + W * (size_t)(j0 - 1));
    cell_load(cur, h, qx, qy, z, n2, colc + W * (size_t)j0);
    cell_derive(prev, hextra);
    // rhs*0 / rhs*1 of the previous row's cell (h, qx, qy)
    R p0[3] = {(R)0, (R)0, (R)0}, p1[3] = {(R)0, (R)0, (R)0};
    for (int r = 0; r <= rend; r++) {
        const int j = j0 + r;
        // prefetch row j+1 while row j is solved
        cell_load(nxt, h, qx, qy, z, n2, colc + W * (size_t)min(j + 1, ny + 1));  // unused at r == rend
        cell_derive(cur, hextra);
        R c0[3] = {(R)0, (R)0, (R)0}, c1[3] = {(R)0, (R)0, (R)0};
        if (r < MR) {  // uniform: row j is a strip row
            CellV<R> L;
            L.h = lane_shift<true>(cur.h);
            L.qx = lane_shift<true>(cur.qx);
            L.qy = lane_shift<true>(cur.qy);
            L.z = lane_shift<true>(cur.z);
            L.sq = lane_shift<true>(cur.sq);
            L.u = lane_shift<true>(cur.u);
            L.v = lane_shift<true>(cur.v);
            R d1[3] = {(R)0, (R)0, (R)0}, d2[3] = {(R)0, (R)0, (R)0};
            bool ok = false;
            if (xedge && j <= ny)
                ok = roe<R, 0>(L.h, cur.h, L.qx, L.qy, L.z, cur.qx, cur.qy, cur.z, L.sq, cur.sq, L.u, L.v, cur.u,
                               cur.v, dx, dt, d1, d2);
            // side 2 of the edge on its left: rhs*1 = (triton_appfis.hpp:821-823)
            if (ok) {
#pragma unroll
                for (int k = 0; k < 3; k++) c1[k] = scale(d2[k]);
            }
            // side 1 of the edge on its right (solved by lane l+1): rhs*0 = (817-819)
            const int okr = from_right(ok ? 1 : 0);
            R d1r[3];
#pragma unroll
            for (int k = 0; k < 3; k++) d1r[k] = lane_shift<false>(d1[k]);
            if (okr && lane < 63) {
#pragma unroll
                for (int k = 0; k < 3; k++) c0[k] = scale(d1r[k]);
            }
        }
        {
            // Y edge (j, j-1): row j is side 1 (rhs*0 +=, 1034-1036), row j-1 side 2
            // (rhs*1 +=, 1038-1040)
            R d1[3] = {(R)0, (R)0, (R)0}, d2[3] = {(R)0, (R)0, (R)0};
            bool ok = false;
            if (cell)
                ok = roe<R, 1>(cur.h, prev.h, cur.qx, cur.qy, cur.z, prev.qx, prev.qy, prev.z, cur.sq, prev.sq, cur.u,
                               cur.v, prev.u, prev.v, dx, dt, d1, d2);
            if (ok) {
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    c0[k] = c0[k] + scale(d1[k]);
                    p1[k] = p1[k] + scale(d2[k]);
                }
            }
        }
        // row j-1 is complete: update with implicit Manning friction (1046-1092) and rain
        if (r >= 1 && cell && j - 1 <= ny) {
            const size_t c = (size_t)col + W * (size_t)(j - 1);
            const R hn = prev.h, qxc = prev.qx, qyc = prev.qy;
            R hij, qxij, qyij;
            hij = hn - p0[0] - p1[0];
            if (hij < (R)0.0) hij = (R)0.0;
            if (hij < hextra) {
                qxij = (R)0.0;
                qyij = (R)0.0;
            } else {
                const R mx = qxc - p0[1] - p1[1];
                const R my = qyc - p0[2] - p1[2];
                const R modM = fsqrt(mx * mx + my * my);
                const R nn = prev.n2;
                if (nn > EPS12 && hn >= hextra && modM > EPS12) {
                    const R tt = fdiv(dt * G * nn * modM, hn * hn * fcbrt(hn));
                    qxij = fdiv((R)-0.5 * (mx - mx * fsqrt((R)1.0 + (R)4.0 * tt)), tt);
                    qyij = fdiv((R)-0.5 * (my - my * fsqrt((R)1.0 + (R)4.0 * tt)), tt);
                } else {
                    qxij = mx;
                    qyij = my;
                }
            }
            hout[c] = hij + rain * dt;
            qxout[c] = qxij;
            qyout[c] = qyij;
        }
        prev = cur;
        cur = nxt;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            p0[k] = c0[k];
            p1[k] = c1[k];
        }
    }
}

// Point sources of the step (triton_appfis.hpp:358-377 on the host, then the functor of
// 1115-1141), one block per scenario, after K1 wrote the updated depth and before K2:
// thread 0 advances the hydrograph interval of the step's start time from the row the
// previous step stopped at; thread k then interpolates source k's flow linearly in time
// and adds flow*dt/dx^2 to its cell (the cells are distinct).
template <typename R>
__global__ __launch_bounds__(64) void k_sources(const Scen<R>* __restrict__ sc, int par_out, Params P)
{
    const Scen<R>& S = sc[blockIdx.x];
    if (S.nsrc == 0 || finished(S.st, P)) return;
    __shared__ int lohi[2];
    ScenState* st = S.st;
    const double simtime = st->simtime, gdt = st->dt;
    if (threadIdx.x == 0) {
        int lo = st->src_low, hi = S.src_rows - 1;
        bool found = false;
        for (int i = lo; i < hi; i++)
            if (S.src_time[i + 1] > simtime) {
                lo = i;
                hi = i + 1;
                found = true;
                break;
            }
        if (!found) lo = hi;
        st->src_low = lo;
        lohi[0] = lo;
        lohi[1] = hi;
    }
    __syncthreads();
    const int lo = lohi[0], hi = lohi[1];
    R* h = S.h[par_out];
    for (int k = threadIdx.x; k < S.nsrc; k += blockDim.x) {
        const double fl = S.src_flow[(size_t)lo * S.nsrc + k], fh = S.src_flow[(size_t)hi * S.nsrc + k];
        const double tl = S.src_time[lo], th = S.src_time[hi];
        double flow;
        if (lo == hi) flow = fl;
        else flow = fl + (((fh - fl) * (simtime - tl)) / (th - tl));
        const int c = S.src_cell[k];
        double hij = (double)h[c];
        hij += (flow * gdt) / (P.dx * P.dx);
        h[c] = (R)hij;
    }
}

__device__ inline void two_sum(double& s, double& cmp, double x)
{
    double t = __dadd_rn(s, x);
    double bp = __dsub_rn(t, s);
    cmp = __dadd_rn(cmp, __dadd_rn(__dsub_rn(s, __dsub_rn(t, bp)), __dsub_rn(x, bp)));
    s = t;
}

// A compensated partial (sum, running error) and the wave reductions of a CFL minimum
// and such a partial (DPP tree; the result is in lane WAVE_TOTAL_LANE)
struct VolSum {
    double s, c;
};
__device__ inline VolSum vol_merge(VolSum a, VolSum b)
{
    two_sum(a.s, a.c, b.s);
    a.c = __dadd_rn(a.c, b.c);
    return a;
}
template <typename R>
__device__ inline void wave_min_sum(R& m, double& s, double& c)
{
    m = wave_reduce(m, RealLim<R>::max(), [](R a, R b) { return b < a ? b : a; });
    const VolSum v = wave_reduce(VolSum{s, c}, VolSum{0.0, 0.0}, vol_merge);
    s = v.s;
    c = v.c;
}

// K2 (wet/dry fix + CFL + volume) as a row march over the same strips as K1: a wave
// walks the cell rows of its strip holding rows j-1, j, j+1 of h and z in registers,
// with the left / right neighbours' values shifted in across lanes, so each cell costs
// one load of h, z, qx and qy.  The momentum fix of each cell is written in place (it
// reads only h and z, which nothing here writes).  Each lane folds its cells into its
// own CFL minimum and compensated volume sum, and the wave's DPP reduction leaves one
// partial per strip, partials[3t .. 3t+2] = (min, sum, compensation).  No LDS and no
// barriers.  MODE 0: initial CFL only (no fix); MODE 1: end of a step.
template <typename R, int MODE>
__global__ __launch_bounds__(BT) void k_fix_march(const Scen<R>* __restrict__ sc, int nscen, int nstrips, int par,
                                                  Params P)
{
    const int lane = threadIdx.x & 63;
    const int strip = blockIdx.x * MWAVES + (threadIdx.x >> 6);  // uniform per wave
    if (strip >= nstrips) return;
    const int s = find_scen(sc, nscen, strip);
    const Scen<R>& S = sc[s];
    if (MODE == 1 && finished(S.st, P)) return;
    const int t = strip - S.strip0;
    const int ty = t / S.strips_x, tx = t - ty * S.strips_x;
    const int nx = S.nx, ny = S.ny;
    const size_t W = (size_t)nx + 2;
    const int col = tx * MW + lane;  // padded column of this lane
    const int j0 = ty * MR + 1;
    const bool colmem = col <= nx + 1;
    const bool cell = lane >= 1 && lane <= MW && col <= nx;
    const int nr = min(MR, ny + 1 - j0);  // cell rows of the strip (j0 <= ny)
    const gmem<R>* h = to_global((const R*)S.h[par]);
    gmem<R>* qx = to_global(S.qx[par]);
    gmem<R>* qy = to_global(S.qy[par]);
    const gmem<R>* z = to_global(S.z);
    const R G = (R)9.81, EPS12 = (R)1e-12;
    const R hextra = (R)P.hextra, dx = (R)P.dx, cn = (R)P.cn;
    auto ld = [&](const gmem<R>* a, int row) { return colmem ? a[(size_t)col + W * (size_t)row] : (R)0.0; };
    R cfl = RealLim<R>::max();
    double vs = 0.0, vc = 0.0;
    R hm = ld(h, j0 - 1), hc = ld(h, j0), zm = (R)0.0, zc = (R)0.0;
    if (MODE == 1) {
        zm = ld(z, j0 - 1);
        zc = ld(z, j0);
    }
    for (int r = 0; r < nr; r++) {
        const int j = j0 + r;
        const R hp = ld(h, j + 1);  // j + 1 <= ny + 1
        const R zp = MODE == 1 ? ld(z, j + 1) : (R)0.0;
        R qxv = cell ? qx[(size_t)col + W * (size_t)j] : (R)0.0;
        R qyv = cell ? qy[(size_t)col + W * (size_t)j] : (R)0.0;
        R hl = (R)0.0, hr = (R)0.0, zl = (R)0.0, zr = (R)0.0;
        if (MODE == 1) {
            hl = lane_shift<true>(hc);
            hr = lane_shift<false>(hc);
            zl = lane_shift<true>(zc);
            zr = lane_shift<false>(zc);
        }
        if (cell) {
            const size_t c = (size_t)col + W * (size_t)j;
            const R hij = hc;
            if (MODE == 1 && hij > hextra) {  // triton_appfis.hpp:1095-1112
                const R zij = zc;
                if (((hij + zij < zr) && (hr < EPS12)) || ((hij + zij < zl) && (hl < EPS12))) {
                    qxv = (R)0.0;
                    qx[c] = qxv;
                }
                if (((hij + zij < zm) && (hm < EPS12)) || ((hij + zij < zp) && (hp < EPS12))) {
                    qyv = (R)0.0;
                    qy[c] = qyv;
                }
            }
            if (hij > hextra) {  // triton_appfis.hpp:567-586
                const R dtx = fdiv(dx, fabs(fdiv(qxv, hij)) + fsqrt(G * hij));
                const R dty = fdiv(dx, fabs(fdiv(qyv, hij)) + fsqrt(G * hij));
                const R cc = dtx < dty ? cn * dtx : cn * dty;
                cfl = cc < cfl ? cc : cfl;
            }
            two_sum(vs, vc, (double)hij);
        }
        hm = hc;
        hc = hp;
        zm = zc;
        zc = zp;
    }
    wave_min_sum(cfl, vs, vc);
    if (lane == WAVE_TOTAL_LANE) {
        double* pp = S.partials + 3 * (size_t)t;
        pp[0] = (double)cfl;
        pp[1] = vs;
        pp[2] = vc;
    }
}

// Finalise: one block per scenario combines its tile partials in a fixed order (the
// reference's Execute2DReduce(MIN) plus the config's FP64 Kahan volume), then does the
// host-side bookkeeping of triton_appfis.hpp:322-335, 381-385 on the device.
template <typename R, int MODE>
__global__ __launch_bounds__(BT) void k_finalize(const Scen<R>* __restrict__ sc, Params P)
{
    __shared__ double smin[BT / 64];
    __shared__ double ssum[BT / 64][2];
    const Scen<R>& S = sc[blockIdx.x];
    if (MODE == 1 && finished(S.st, P)) return;  // before thread 0 updates simtime below
    double m = (double)RealLim<R>::max();
    double as = 0.0, ac = 0.0;
    for (int p = threadIdx.x; p < S.nstrips; p += BT) {
        const double* pp = S.partials + 3 * (size_t)p;
        m = pp[0] < m ? pp[0] : m;
        two_sum(as, ac, pp[1]);
        ac = __dadd_rn(ac, pp[2]);
    }
    wave_min_sum(m, as, ac);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == WAVE_TOTAL_LANE) {
        smin[wid] = m;
        ssum[wid][0] = as;
        ssum[wid][1] = ac;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double bs = ssum[0][0], bc = ssum[0][1];
    for (int w = 1; w < BT / 64; w++) {
        m = smin[w] < m ? smin[w] : m;
        two_sum(bs, bc, ssum[w][0]);
        bc = __dadd_rn(bc, ssum[w][1]);
    }
    m = smin[0] < m ? smin[0] : m;
    ScenState* st = S.st;
    const double PI = P.print_interval;
    double simtime = st->simtime;
    int pid = st->print_id;
    if (MODE == 1) {
        const double dtn = st->dt;
        simtime += dtn;
        if (PI > 0.0 && simtime >= PI * (pid + 1)) pid++;
        const int stp = st->step;
        if (stp < P.nsteps_rec) {
            const double vol = __dmul_rn(__dmul_rn(__dadd_rn(bs, bc), P.dx), P.dx);
            if (S.rec_time) S.rec_time[stp] = simtime;
            if (S.rec_vol) S.rec_vol[stp] = vol;
            if (S.rec_dt) S.rec_dt[stp] = dtn;
        }
        st->step = stp + 1;
        st->simtime = simtime;
        st->print_id = pid;
    }
    // next dt (triton_appfis.hpp:322-335): the minimum is an R value held exactly in m;
    // with time_increment_fixed the reference skips that block and keeps time_step
    double gdt = P.dt_fallback;
    if (!P.fixed_dt) {
        gdt = (double)(R)m;
        if (gdt >= (double)RealLim<R>::max() - 1.0) gdt = P.dt_fallback;
        if (PI > 0.0 && simtime + gdt > PI * (pid + 1)) gdt = PI * (pid + 1) - simtime;
    }
    st->dt = (double)(R)gdt;
}

__global__ void k_init_state(ScenState* st, int n, double start_time)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    st[s].dt = 0.0;
    st[s].simtime = start_time;
    st[s].print_id = 0;
    st[s].step = 0;
    st[s].src_low = 0;
}

template <typename R>
int run_flood(appfis_ctx* ctx, int nscen, const int* nx, const int* ny, void* const* h, void* const* qx,
              void* const* qy, const void* const* z, const void* const* n2, const appfis_flood_params* prm, int steps,
              bool graph, double* time_host, double* volume_host, double* dt_host,
              const appfis_flood_sources* src = nullptr, int* steps_done = nullptr)
{
    APPFIS_NOT_CAPTURING(ctx, "the flood step loop (it synchronises and captures graphs of its own)");
    const int nsrc = src ? src->count : 0;
    if (nsrc > 0) {
        APPFIS_REQUIRE(nscen == 1, "point sources are supported for one scenario per call");
        APPFIS_REQUIRE(src->cell && src->time && src->flow && src->rows >= 1, "sources: NULL table or no rows");
        std::vector<int> seen(src->cell, src->cell + nsrc);
        std::sort(seen.begin(), seen.end());
        for (int k = 0; k < nsrc; k++) {
            const int c = src->cell[k], W = nx[0] + 2, i = c % W, j = c / W;
            APPFIS_REQUIRE(c >= 0 && i >= 1 && i <= nx[0] && j >= 1 && j <= ny[0],
                           "source %d: cell %d is not an interior cell", k, c);
            APPFIS_REQUIRE(k == 0 || seen[k] != seen[k - 1], "sources: cell %d appears twice", seen[k]);
        }
    }
    for (int s = 0; s < nscen; s++) {
        APPFIS_REQUIRE(nx[s] >= 1 && ny[s] >= 1, "scenario %d: empty interior", s);
        APPFIS_REQUIRE((double)(nx[s] + 2) * (ny[s] + 2) < 2147483647.0, "scenario %d exceeds the int index range", s);
        APPFIS_REQUIRE(h[s] && qx[s] && qy[s] && z[s] && n2[s], "scenario %d: NULL field", s);
    }
    APPFIS_REQUIRE(prm->dx > 0.0, "dx must be positive");
    // strip work-list
    std::vector<Scen<R>> hs((size_t)nscen);
    long long total_strips = 0;
    size_t field_bytes = 0;
    for (int s = 0; s < nscen; s++) {
        hs[s].nx = nx[s];
        hs[s].ny = ny[s];
        const int sx = (int)ceil_div(nx[s], MW), sy = (int)ceil_div(ny[s], MR);
        hs[s].strips_x = sx;
        hs[s].strip0 = (int)total_strips;
        hs[s].nstrips = sx * sy;
        total_strips += (long long)sx * sy;
        field_bytes += ((size_t)(nx[s] + 2) * (ny[s] + 2) * sizeof(R) + 255) / 256 * 256;
    }
    APPFIS_REQUIRE(total_strips < (1LL << 31), "too many strips");
    // per-step records only when the caller asked for them
    const bool want_rec = time_host || volume_host || dt_host;
    const int rec = want_rec && steps > 0 ? steps : 1;
    // workspace: second buffers (h, qx, qy) | states | partials | records | descriptor table
    const size_t st_bytes = ((size_t)nscen * sizeof(ScenState) + 255) / 256 * 256;
    const size_t part_bytes = ((size_t)total_strips * 24 + 255) / 256 * 256;
    const size_t rec_bytes = ((size_t)nscen * rec * 8 + 255) / 256 * 256;
    const size_t tab_bytes = ((size_t)nscen * sizeof(Scen<R>) + 255) / 256 * 256;
    const size_t src_rows = nsrc > 0 ? (size_t)src->rows : 0;
    const size_t src_bytes = ((size_t)nsrc * 4 + 255) / 256 * 256 + (src_rows * 8 + 255) / 256 * 256 +
                             (src_rows * nsrc * 8 + 255) / 256 * 256;
    int stt = ensure_work(ctx, 3 * field_bytes + st_bytes + part_bytes + 3 * rec_bytes + tab_bytes + src_bytes);
    if (stt) return stt;
    char* w = (char*)ctx->work;
    size_t off = 0;
    for (int s = 0; s < nscen; s++) {
        const size_t fb = (size_t)(nx[s] + 2) * (ny[s] + 2) * sizeof(R);
        const size_t fa = (fb + 255) / 256 * 256;
        hs[s].h[0] = (R*)h[s];
        hs[s].qx[0] = (R*)qx[s];
        hs[s].qy[0] = (R*)qy[s];
        hs[s].h[1] = (R*)(w + off);
        hs[s].qx[1] = (R*)(w + off + field_bytes);
        hs[s].qy[1] = (R*)(w + off + 2 * field_bytes);
        off += fa;
        // the second buffers start as full copies (ghost cells included)
        APPFIS_HIP_CHECK(hipMemcpyAsync(hs[s].h[1], h[s], fb, hipMemcpyDeviceToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemcpyAsync(hs[s].qx[1], qx[s], fb, hipMemcpyDeviceToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemcpyAsync(hs[s].qy[1], qy[s], fb, hipMemcpyDeviceToDevice, ctx->stream));
        hs[s].z = (const R*)z[s];
        hs[s].n2 = (const R*)n2[s];
    }
    char* base = w + 3 * field_bytes;
    ScenState* st = (ScenState*)base;
    double* partials = (double*)(base + st_bytes);
    double* rt = (double*)(base + st_bytes + part_bytes);
    double* rv = (double*)(base + st_bytes + part_bytes + rec_bytes);
    double* rd = (double*)(base + st_bytes + part_bytes + 2 * rec_bytes);
    Scen<R>* tab = (Scen<R>*)(base + st_bytes + part_bytes + 3 * rec_bytes);
    char* sbase = base + st_bytes + part_bytes + 3 * rec_bytes + tab_bytes;
    int* d_cell = (int*)sbase;
    double* d_time = (double*)(sbase + ((size_t)nsrc * 4 + 255) / 256 * 256);
    double* d_flow = (double*)((char*)d_time + (src_rows * 8 + 255) / 256 * 256);
    for (int s = 0; s < nscen; s++) {
        hs[s].st = st + s;
        hs[s].partials = partials + 3 * (size_t)hs[s].strip0;
        hs[s].rec_time = rt + (size_t)s * rec;
        hs[s].rec_vol = rv + (size_t)s * rec;
        hs[s].rec_dt = rd + (size_t)s * rec;
        hs[s].nsrc = 0;
        hs[s].src_rows = 0;
        hs[s].src_cell = nullptr;
        hs[s].src_time = nullptr;
        hs[s].src_flow = nullptr;
    }
    if (nsrc > 0) {
        hs[0].nsrc = nsrc;
        hs[0].src_rows = src->rows;
        hs[0].src_cell = d_cell;
        hs[0].src_time = d_time;
        hs[0].src_flow = d_flow;
        APPFIS_HIP_CHECK(hipMemcpyAsync(d_cell, src->cell, (size_t)nsrc * 4, hipMemcpyHostToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemcpyAsync(d_time, src->time, src_rows * 8, hipMemcpyHostToDevice, ctx->stream));
        APPFIS_HIP_CHECK(
            hipMemcpyAsync(d_flow, src->flow, src_rows * nsrc * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    APPFIS_HIP_CHECK(hipMemcpyAsync(tab, hs.data(), (size_t)nscen * sizeof(Scen<R>), hipMemcpyHostToDevice, ctx->stream));
    APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // the host table is a local
    Params P{prm->dx,        prm->courant,   prm->hextra, prm->dt_fallback,
             prm->print_interval, prm->rain_rate, want_rec ? rec : 0, prm->time_increment_fixed ? 1 : 0,
             prm->end_time};
    // with an end time, the host looks at the scenarios' simulated time every
    // CHECK_EVERY steps and stops issuing steps once all of them have reached it (steps
    // issued past that point do nothing on the device)
    const bool until = prm->end_time > 0.0;
    constexpr int CHECK_EVERY = 10 * CHUNK;
    std::vector<ScenState> hst((size_t)nscen);
    auto read_states = [&]() -> int {
        APPFIS_HIP_CHECK(
            hipMemcpyAsync(hst.data(), st, (size_t)nscen * sizeof(ScenState), hipMemcpyDeviceToHost, ctx->stream));
        APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return APPFIS_OK;
    };
    auto all_finished = [&]() {
        for (const ScenState& x : hst)
            if (!(x.simtime >= prm->end_time)) return false;
        return true;
    };
    hipLaunchKernelGGL(k_init_state, dim3(ceil_div(nscen, 64)), dim3(64), 0, ctx->stream, st, nscen, prm->start_time);
    APPFIS_LAUNCH_CHECK();
    const dim3 grid((unsigned)ceil_div(total_strips, MWAVES));  // MWAVES strips per block
    const int nstrips = (int)total_strips;
    // dt of the first step from the initial state
    hipLaunchKernelGGL((k_fix_march<R, 0>), grid, dim3(BT), 0, ctx->stream, tab, nscen, nstrips, 0, P);
    APPFIS_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_finalize<R, 0>), dim3(nscen), dim3(BT), 0, ctx->stream, tab, P);
    APPFIS_LAUNCH_CHECK();
    auto step_pair = [&](int par) -> int {
        hipLaunchKernelGGL((k_flux_march<R>), grid, dim3(BT), 0, ctx->stream, tab, nscen, nstrips, par, P);
        APPFIS_LAUNCH_CHECK();
        if (nsrc > 0) {
            hipLaunchKernelGGL((k_sources<R>), dim3(nscen), dim3(64), 0, ctx->stream, tab, par ^ 1, P);
            APPFIS_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL((k_fix_march<R, 1>), grid, dim3(BT), 0, ctx->stream, tab, nscen, nstrips, par ^ 1, P);
        APPFIS_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_finalize<R, 1>), dim3(nscen), dim3(BT), 0, ctx->stream, tab, P);
        APPFIS_LAUNCH_CHECK();
        return APPFIS_OK;
    };
    int done = 0;
    if (graph && steps >= 2 * CHUNK) {
        // CHUNK steps as one graph, captured on the first call with these arguments and
        // replayed by later ones (the kernels read the scenarios from the table)
        const std::vector<uint64_t> key = {2,
                                           sizeof(R),
                                           key_ptr(tab),
                                           (uint64_t)nscen,
                                           (uint64_t)nstrips,
                                           (uint64_t)(nsrc > 0),
                                           key_bits(P.dx),
                                           key_bits(P.cn),
                                           key_bits(P.hextra),
                                           key_bits(P.dt_fallback),
                                           key_bits(P.print_interval),
                                           key_bits(P.rain),
                                           (uint64_t)P.nsteps_rec,
                                           (uint64_t)P.fixed_dt,
                                           key_bits(P.end_time)};
        hipGraphExec_t ge = nullptr;
        int cst = cached_graph(ctx, key,
                               [&]() {
                                   int s2 = APPFIS_OK;
                                   for (int k = 0; k < CHUNK && s2 == APPFIS_OK; k++) s2 = step_pair(k & 1);
                                   return s2;
                               },
                               &ge);
        if (cst) return cst;
        const int reps = steps / CHUNK;
        bool stop = false;
        for (int r = 0; r < reps && !stop; r++) {
            APPFIS_HIP_CHECK(hipGraphLaunch(ge, ctx->stream));
            if (until && (r + 1) * CHUNK % CHECK_EVERY == 0) {
                cst = read_states();
                if (cst) return cst;
                stop = all_finished();
            }
            done = (r + 1) * CHUNK;
        }
        if (stop) done = steps;
    }
    for (int k = done; k < steps; k++) {
        int cst = step_pair(k & 1);
        if (cst) return cst;
        if (until && (k + 1) % CHECK_EVERY == 0) {
            cst = read_states();
            if (cst) return cst;
            if (all_finished()) break;
        }
    }
    // each scenario's result is in buffer (steps it ran) % 2: bring it home when that is
    // the work copy
    int st_read = read_states();
    if (st_read) return st_read;
    for (int s = 0; s < nscen; s++) {
        if (!(hst[s].step & 1)) continue;
        const size_t fb = (size_t)(nx[s] + 2) * (ny[s] + 2) * sizeof(R);
        APPFIS_HIP_CHECK(hipMemcpyAsync(h[s], hs[s].h[1], fb, hipMemcpyDeviceToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemcpyAsync(qx[s], hs[s].qx[1], fb, hipMemcpyDeviceToDevice, ctx->stream));
        APPFIS_HIP_CHECK(hipMemcpyAsync(qy[s], hs[s].qy[1], fb, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (steps_done) *steps_done = hst[0].step;
    if (want_rec && steps > 0) {
        const size_t rb = (size_t)nscen * rec * 8;
        if (time_host) APPFIS_HIP_CHECK(hipMemcpyAsync(time_host, rt, rb, hipMemcpyDeviceToHost, ctx->stream));
        if (volume_host) APPFIS_HIP_CHECK(hipMemcpyAsync(volume_host, rv, rb, hipMemcpyDeviceToHost, ctx->stream));
        if (dt_host) APPFIS_HIP_CHECK(hipMemcpyAsync(dt_host, rd, rb, hipMemcpyDeviceToHost, ctx->stream));
        if (time_host || volume_host || dt_host) APPFIS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    }
    return APPFIS_OK;
}

// Synthetic DEM (see include/appfis_hip.h): one thread per padded cell.
template <typename R>
__global__ void k_dem(R* z, int nx, int ny, double dx, const double* hills, int nhills, double slope, double wall)
{
    const int W = nx + 2, H = ny + 2;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)W * H) return;
    const int j = (int)(t / (size_t)W), i = (int)(t - (size_t)j * W);
    if (i == 0 || j == 0 || i == W - 1 || j == H - 1) {
        z[t] = (R)wall;
        return;
    }
    const double x = ((double)i - 0.5) * dx, y = ((double)j - 0.5) * dx;
    double v = 0.0;
    for (int k = 0; k < nhills; k++) {
        const double ddx = x - hills[4 * k + 2], ddy = y - hills[4 * k + 3], s = hills[4 * k + 1];
        v = v + hills[4 * k] * exp(-(ddx * ddx + ddy * ddy) / (2.0 * s * s));
    }
    z[t] = (R)(v + slope * x);
}

}  // namespace

namespace appfis {
// run_flood<float>, defined in flood_f32.hip (FMA-contracted translation unit)
int flood_run_f32(appfis_ctx* ctx, int nscen, const int* nx, const int* ny, void* const* h, void* const* qx,
                  void* const* qy, const void* const* z, const void* const* n2, const appfis_flood_params* prm,
                  int steps, bool graph, double* time_host, double* volume_host, double* dt_host,
                  const appfis_flood_sources* src, int* steps_done);
}  // namespace appfis
