// tt06_model.hpp -- TT06Dev model data shared by host and device code.
//
// State layout, constants, fit layout and cell-type parameters of the
// TT06Dev reaction (configuration B: rational fits at tolerance 1e-4 with
// mod = 1, "v2" gate kernels, generic non-gate update).
//
// References (reference tree):
//   state order           elec/TT06Func.h:10-14
//   constants c2..c44, f* elec/TT06NonGates.c:11-251
//   fit order             elec/TT06Dev_Reaction.cc:206-207, :218-219
//   fit orders (l, m)     hard-coded in elec/TT06GatesSimd.c (per gate), and
//                         those of the fit file test/tests/multipleReaction/fit.data
//   cell types            elec/TT06Func.cc:421-612
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdh {

enum StateSlot : int {
   S_CA_I = 0, S_DVK_I, S_NA_I, S_CA_SS, S_CA_SR, S_R_PRIME, S_FCASS,
   S_M, S_H, S_J, S_XR1, S_XR2, S_XS, S_R, S_D, S_F, S_F2, S_JL, S_S,
   N_STATE
};

// Non-gate constants, derived on the host from the TT06 physical parameters.
struct TT06Const {
   double c2, c3, c4, c5, c6, c7, c8, c9, c11, c12, c13, c14, c15, c16, c17,
          c18, c19, c20, c21, c22, c23, c24, c25, c26, c27, c28, c29, c30,
          c31, c32, c33, c34, c36, c40, c43, c44;
   double exp_coef[17];   // 1/i!, the m-gate Rush-Larsen series
};

// Configuration A (tolerance 0, the exact TT06 functions): the constants of
// fv0..fv5 (elec/TT06NonGates.c:236-248) and, for mod = 1, the cubic
// Hermite bridges of the h and j rates, {x0, x1, c0, c1, c2, c3} with
// 1/tau = c0 + x*(c1 + x*(c2 + x*c3)), x = V - x0, on [x0, x1]
// (elec/TT06Tau.cc:131-157).
struct TT06ExactConst {
   double f1, f2, f3, f4, f5, f6, f7a, f9a, f10;
   double hmod[6], jmod[6];
   int mod;
   int pad_;
};

// Fit orders (l, m) in fit order; the kernels are specialised on them.
struct FitOrder { int l, m; };
constexpr FitOrder kFitOrder[33] = {
   {1, 10}, {1, 6}, {1, 8}, {1, 11}, {1, 9}, {5, 8}, {11, 7},          // fv0..fv6
   {10, 5}, {1, 18},   // m
   {7, 9}, {11, 11},   // h
   {7, 9}, {1, 13},    // j
   {5, 8}, {1, 13},    // Xr1
   {1, 10}, {1, 10},   // Xr2
   {5, 5}, {6, 9},     // Xs
   {5, 8}, {5, 7},     // r
   {5, 7}, {7, 10},    // d
   {5, 8}, {13, 11},   // f
   {8, 5}, {11, 12},   // f2
   {6, 12}, {1, 1},    // jL
   {7, 8}, {6, 7},     // s0
   {7, 8}, {11, 11},   // s1
};

// Device layout of the coefficients: fit i starts at fit_offset(i), with one
// unused slot between consecutive fits.  The gap keeps the compiler from
// merging the scalar loads of different fits into one wide load issued at
// kernel entry (which would hold all 468 coefficients in SGPRs at once).
constexpr int kFitGap = 1;
constexpr int fit_offset(int i)
{
   int o = 0;
   for (int k = 0; k < i; k++) o += kFitOrder[k].l + kFitOrder[k].m + kFitGap;
   return o;
}
constexpr int kFitCoefCount = 468;                // coefficients in a fit file
constexpr int kFitCoefTotal = fit_offset(33);     // device slots incl. gaps

// Engine layout of per-cell arrays with several fields (states, weights):
// blocks of 64 cells (one wavefront), field-major inside a block, so that a
// wavefront's fields are one contiguous run of memory.  Field q of cell i of
// an array with nq fields sits at blk_base(i, nq) + 64*q.
constexpr int kBlk = 64;
__host__ __device__ inline int64_t blk_base(int64_t i, int nq)
{
   return (i >> 6) * (int64_t)nq * kBlk + (i & (kBlk - 1));
}
__host__ __device__ inline int64_t blk_index(int64_t q, int64_t i, int nq)
{
   return blk_base(i, nq) + (int64_t)kBlk * q;
}

// vector types for loads of several fields per lane
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The drop-in stencil's 18 weights per cell, in 64-cell tiles with the
// fields in groups that one load instruction each brings in: fields 0-3,
// 4-7, 8-11 and 12-15 as 16 bytes per lane, fields 16-17 as 8 bytes per
// lane.  A tile's groups are one contiguous 4.5 KB run, as in blk_index.
__host__ __device__ inline int64_t grp_index18(int k, int64_t i)
{
   const int64_t base = (i >> 6) * (18 * kBlk);
   const int64_t lane = i & (kBlk - 1);
   return k < 16 ? base + (k >> 2) * (4 * kBlk) + 4 * lane + (k & 3)
                 : base + 16 * kBlk + 2 * lane + (k - 16);
}

// The mixed engine's 9 weight words per cell the same way: words 0-3 and
// 4-7 as 16 bytes per lane, word 8 as 4 bytes per lane.
__host__ __device__ inline int64_t grp_index9(int k, int64_t i)
{
   const int64_t base = (i >> 6) * (9 * kBlk);
   const int64_t lane = i & (kBlk - 1);
   return k < 8 ? base + (k >> 2) * (4 * kBlk) + 4 * lane + (k & 3) : base + 8 * kBlk + lane;
}

// The mixed engine's state arrays: per slab, runs of kRun cells, field-major
// inside a run (field q of the run's cells one contiguous stretch), the last
// run as wide as the cells left (a multiple of 64, as slabs are).  Measured
// on the 20.1M-cell shell on one MI355X, in one process: the mixed step
// takes 0.922 ms with 4096-cell runs against 0.930 ms with 64-cell blocks
// (0.928 with 1024 or 2048, 0.932 with 8192, 0.949 with 16384): the width
// changes how the concurrent read-and-write streams fall on the HBM channels,
// measured rather than modelled.
constexpr int64_t kRun = 4096;
__host__ __device__ inline int64_t run_width(int64_t k, int64_t S)
{
   const int64_t r0 = k & ~(kRun - 1);
   return S - r0 < kRun ? S - r0 : kRun;
}
// field q of slab cell k of an nq-field array over a slab of S cells
__host__ __device__ inline int64_t run_index(int64_t q, int64_t k, int nq, int64_t S)
{
   const int64_t r0 = k & ~(kRun - 1);
   return r0 * nq + q * run_width(k, S) + (k - r0);
}

// The FP64 engine's states: per slab, runs of kRun cells as run_index, with
// the fields in pairs: fields 2p and 2p + 1 of a cell side by side (16
// bytes per lane), the pairs one after the other, and with nq odd the last
// field alone after them.  On one MI355X the FP64 FMA step takes 1.432 ms
// on the 20.1 M-cell shell this way against 1.457 ms with 64-cell blocks
// and one field per access (reference rounding 1.558 against 1.563), bit
// for bit the same; runs with one field per access had measured 1.460
// against 1.446 ms.  The mixed engine's fields measured slower in pairs
// (its FP32 fields 0.976 against 0.908 ms, its two FP64 ones alone 0.908
// against 0.902 ms) and keep run_index.
__host__ __device__ inline int64_t run_pair_index(int64_t q, int64_t k, int nq, int64_t S)
{
   const int64_t r0 = k & ~(kRun - 1);
   const int64_t w = run_width(k, S), j = k - r0;
   return q < (nq & ~1) ? r0 * nq + (q & ~1) * w + 2 * j + (q & 1) : r0 * nq + q * w + j;
}

// Per-cell-type parameters (a reaction object's CellTypeParms).
struct CellType {
   int s_switch;
   double P_NaK, g_Ks, g_Kr, g_to, g_NaL;
   double Vm0;
   double state0[N_STATE];   // K_i (not dVK_i) in slot S_DVK_I
};

constexpr int kMaxCellTypes = 8;

// per-cell flag bits (see tt06_f64.hip): the reference's v2 gate loops stop
// at 4*floor(n/4) of a reaction group, and its f-gate main loop updates
// lanes 0..3 of each 12-cell group with a different rounding order.
constexpr uint8_t kFlagGatesFixed = 1;
constexpr uint8_t kFlagFGateSub = 2;

inline uint8_t gate_flags(int64_t i, int64_t n)
{
   uint8_t f = 0;
   if (i >= 4 * (n / 4)) f |= kFlagGatesFixed;
   if (i < 12 * (n / 12) && (i % 12) < 4) f |= kFlagFGateSub;
   return f;
}

// host-side helpers (tt06_host.cpp)
void tt06_constants(TT06Const *k);
void tt06_exact_constants(TT06ExactConst *e, int mod);
bool tt06_standard_celltype(const char *name, CellType *out);

}  // namespace cdh
