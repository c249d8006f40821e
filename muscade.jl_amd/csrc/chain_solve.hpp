// On-device solve of the assembled system Lλx·Δx = Lλ for chain topologies, and the
// Newton-Newmark decrement that consumes Δx (the host steps of src/SweepX.jl:205-211 and
// src/SweepX.jl:98-132 for models whose elements join their nodes into paths).
//
// Block layout
//   The X dofs of one node form one block of B (3 or 6) positions; blocks follow the
//   nodes along each path, paths one after another. Lλx is then block tridiagonal:
//     D[i]  rows of block i, columns of block i        (B×B, column-major: entry (r, c)
//     L[i]  rows of block i, columns of block i-1       at i·B² + c·B + r, so that the rows
//     U[i]  rows of block i, columns of block i+1       of one CSC column land contiguously)
//   Two paths meet with zero L/U blocks, so many chains form one system. A node with
//   fewer than B dofs is padded with identity rows (zero solution).
//   At prepare, slot_kernel gives each nz of Lλx its destination (slot) in D|L|U, and the
//   slots no nz covers are listed; a solve zeroes those, then scatters nzval by slot with
//   coalesced reads.
//   The λ dof of a Hold (src/BasicElements.jl:426-438) is coupled to its held dof h only:
//   row λ reads cL·x_h = r_λ (its diagonal is zero), row h holds cU·λ. The λ row fixes x_h;
//   the pack replaces row h of the block system by that equation, saving the original row
//   so that λ = (r_h − Σ_j K_hj x_j)/cU once x is known.
//
// Block cyclic reduction (no pivoting between blocks, partial pivoting inside a block)
//   level s = 1, 2, 4, …: blocks i ≡ s (mod 2s) are eliminated, blocks j ≡ 0 (mod 2s) kept.
//   factor:  D_i = P⁻¹LU; L_i ← D_i⁻¹L_i, U_i ← D_i⁻¹U_i, b_i ← D_i⁻¹b_i        (in place)
//   update:  D_j −= L_j U_{j−s} + U_j L_{j+s};  b_j −= L_j b_{j−s} + U_j b_{j+s};
//            L_j ← −L_j L_{j−s} (couples j−2s),  U_j ← −U_j U_{j+s} (couples j+2s)
//   top:     x_0 = D_0⁻¹ b_0
//   back:    for s descending, x_i = b_i − L_i x_{i−s} − U_i x_{i+s}     (x overwrites b)
//   Every eliminated block is written once and read-only afterwards, so the levels run
//   in place without a second copy; a kept block is written only by its own lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace chain {

constexpr int kCB = 256;

enum : unsigned {
    kBadPattern = 1u,    // an entry outside the block-tridiagonal / λ-pair structure
    kSingular = 2u,      // zero or non-finite pivot in a diagonal block
    kBadPair = 4u,       // a λ row with a nonzero diagonal or a zero coupling
};

// LU with partial pivoting of one B×B block held in registers (rows swapped whole, as
// LAPACK getrf): P·A = L·U with unit-lower L below the diagonal of a, U on and above it.
// All indices are compile-time after unrolling; the pivot row is chosen with selects.
template <int B>
__device__ inline bool lu_factor(double (&a)[B][B], int (&piv)[B]) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < B; ++k) {
        int p = k;
        double m = fabs(a[k][k]);
#pragma unroll
        for (int r = k + 1; r < B; ++r) {
            const double v = fabs(a[r][k]);
            if (v > m) { m = v; p = r; }
        }
        piv[k] = p;
#pragma unroll
        for (int r = k + 1; r < B; ++r) {
            const bool sw = (p == r);
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const double t = a[k][j];
                a[k][j] = sw ? a[r][j] : t;
                a[r][j] = sw ? t : a[r][j];
            }
        }
        const double d = a[k][k];
        ok = ok && d != 0.0 && isfinite(d);
#pragma unroll
        for (int r = k + 1; r < B; ++r) {
            const double f = a[r][k] / d;
            a[r][k] = f;
#pragma unroll
            for (int j = k + 1; j < B; ++j) a[r][j] -= f * a[k][j];
        }
    }
    return ok;
}

// v ← A⁻¹v from lu_factor's output: the row swaps in order, then L and U substitution.
template <int B>
__device__ inline void lu_solve(const double (&a)[B][B], const int (&piv)[B], double (&v)[B]) {
#pragma unroll
    for (int k = 0; k < B; ++k) {
#pragma unroll
        for (int r = k + 1; r < B; ++r) {
            const bool sw = (piv[k] == r);
            const double t = v[k];
            v[k] = sw ? v[r] : t;
            v[r] = sw ? t : v[r];
        }
    }
#pragma unroll
    for (int k = 0; k < B; ++k)
#pragma unroll
        for (int r = k + 1; r < B; ++r) v[r] -= a[r][k] * v[k];
#pragma unroll
    for (int k = B - 1; k >= 0; --k) {
        double s = v[k];
#pragma unroll
        for (int j = k + 1; j < B; ++j) s -= a[k][j] * v[j];
        v[k] = s / a[k][k];
    }
}

// The status words of a solve: flag (the k* bits) at word 0, then, as 64-bit values at
// byte 8 and 16, the first block found with a bad pivot (the earliest level; a block
// i = s + 2s·t of level s, or 0 for the last remaining block) and the first bad λ pair
// (all ones: none).
constexpr int kStatusBytes = 24;
__device__ inline unsigned long long* status_where(unsigned* f) { return reinterpret_cast<unsigned long long*>(f + 2); }
__global__ void clear_flag(unsigned* f) {
    f[0] = 0u;
    status_where(f)[0] = ~0ull;
    status_where(f)[1] = ~0ull;
}

__global__ __launch_bounds__(kCB) void fill_f64(int64_t n, double* __restrict__ p, double v) {
    const int64_t i = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kCB) void fill_i32(int64_t n, int32_t* __restrict__ p, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (i < n) p[i] = v;
}

// Identity rows of padding positions (their slots are uncovered, so zeroed before).
template <int B>
__global__ __launch_bounds__(kCB) void dummy_kernel(int64_t n, const int32_t* __restrict__ dpos, double* __restrict__ D,
                                                    double* __restrict__ b) {
    const int64_t q = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (q >= n) return;
    const int64_t p = dpos[q], blk = p / B, i = p - blk * B;
    D[blk * B * B + i * B + i] = 1.0;
    b[p] = 0.0;
}

// Prepare: one thread per column c of the CSC pattern. slot[k] of each nz k: ≥ 0, its index
// in D|L|U; −(4p + w + 1), coupling w (0: cL, row λ col h; 1: cU, row h col λ; 2: the λ
// diagonal) of λ pair p. unc[q] is cleared for every slot q that an nz covers.
// pos[d] ≥ 0: block position of dof d; pos[d] = −(p+1): the λ dof of pair p.
template <int B>
__global__ __launch_bounds__(kCB) void slot_kernel(int64_t nX, int64_t nb, const int32_t* __restrict__ colptr,
                                                   const int32_t* __restrict__ rowval, const int32_t* __restrict__ pos,
                                                   const int32_t* __restrict__ hpos, int32_t* __restrict__ slot,
                                                   int32_t* __restrict__ unc, unsigned* __restrict__ flag) {
    const int64_t c = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (c >= nX) return;
    const int64_t BB = (int64_t)B * B;
    const int32_t pcol = pos[c];
    bool bad = false;
    const int32_t k1 = colptr[c + 1];
    for (int32_t k = colptr[c]; k < k1; ++k) {
        const int32_t prow = pos[rowval[k]];
        int64_t q = -1;
        if (pcol >= 0 && prow >= 0) {
            const int64_t br = prow / B, bc = pcol / B;
            const int64_t o = br * BB + (pcol - bc * B) * B + (prow - br * B);
            if (br == bc) q = o;
            else if (br == bc + 1) q = nb * BB + o;
            else if (br + 1 == bc) q = 2 * nb * BB + o;
            else bad = true;
            if (q >= 0) { slot[k] = (int32_t)q; unc[q] = 0; }
        } else if (pcol < 0 && prow < 0) {
            if (pcol == prow) slot[k] = -(4 * (-pcol - 1) + 2 + 1);
            else bad = true;
        } else if (pcol < 0) {
            const int32_t p = -pcol - 1;
            if (prow == hpos[p]) slot[k] = -(4 * p + 1 + 1);
            else bad = true;
        } else {
            const int32_t p = -prow - 1;
            if (pcol == hpos[p]) slot[k] = -(4 * p + 0 + 1);
            else bad = true;
        }
        if (bad) slot[k] = INT32_MIN;
    }
    if (bad) atomicOr(flag, kBadPattern);
}

// Prepare: the list of uncovered slots (positions from an exclusive scan of unc).
__global__ __launch_bounds__(kCB) void compact_kernel(int64_t n, const int32_t* __restrict__ unc,
                                                      const int32_t* __restrict__ at, int32_t* __restrict__ list) {
    const int64_t q = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (q < n && unc[q]) list[at[q]] = (int32_t)q;
}

__global__ __launch_bounds__(kCB) void zero_slots_kernel(int64_t n, const int32_t* __restrict__ list,
                                                         double* __restrict__ DLU) {
    const int64_t q = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (q < n) DLU[list[q]] = 0.0;
}

// Prepare: slots as offsets into the whole chain buffer D|L|U|b|pc (a λ coupling −(4p+w+1)
// becomes pcbase + 4p + w), so that one map serves the pack and an assembly that scatters
// straight into the blocks.
__global__ __launch_bounds__(kCB) void absolute_slot_kernel(int64_t nnz, int64_t pcbase, int32_t* __restrict__ slot) {
    const int64_t k = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (k >= nnz) return;
    const int32_t q = slot[k];
    if (q < 0 && q != INT32_MIN) slot[k] = (int32_t)(pcbase - (int64_t)q - 1);
}

// One thread per nz: nzval into the chain buffer by slot (coalesced reads; the rows of one
// column are contiguous slots of D|L|U; λ couplings go to pc).
__global__ __launch_bounds__(kCB) void pack_kernel(int64_t nnz, const int32_t* __restrict__ slot,
                                                   const double* __restrict__ nzval, double* __restrict__ cbuf) {
    const int64_t k = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (k >= nnz) return;
    const int32_t q = __builtin_nontemporal_load(slot + k);
    cbuf[q] = __builtin_nontemporal_load(nzval + k);
}

// One thread per dof: Lλ into b, or into the λ pair's r_λ.
__global__ __launch_bounds__(kCB) void pack_rhs_kernel(int64_t nX, const int32_t* __restrict__ pos,
                                                       const double* __restrict__ rhs, double* __restrict__ b,
                                                       double* __restrict__ pc) {
    const int64_t d = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (d >= nX) return;
    const int32_t p = pos[d];
    if (p >= 0) b[p] = rhs[d];
    else pc[4 * (int64_t)(-p - 1) + 3] = rhs[d];
}

// One thread per λ pair: save row h (L | D | U parts and r_h), replace it by x_h = r_λ/cL.
template <int B>
__global__ __launch_bounds__(kCB) void pair_kernel(int64_t np, const int32_t* __restrict__ hpos,
                                                   const double* __restrict__ pc, double* __restrict__ D,
                                                   double* __restrict__ L, double* __restrict__ U,
                                                   double* __restrict__ b, double* __restrict__ saved,
                                                   unsigned* __restrict__ flag) {
    const int64_t p = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (p >= np) return;
    const double cL = pc[4 * p], cU = pc[4 * p + 1], dg = pc[4 * p + 2], rl = pc[4 * p + 3];
    if (dg != 0.0 || cL == 0.0 || cU == 0.0) {
        atomicOr(flag, kBadPair);
        atomicCAS(status_where(flag) + 1, ~0ull, (unsigned long long)p);
    }
    const int64_t h = hpos[p], blk = h / B, i = h - blk * B;
    double* sv = saved + p * (3 * B + 1);
    const int64_t o = blk * B * B + i;    // row i of the blocks: entries o + j·B
#pragma unroll
    for (int j = 0; j < B; ++j) {
        sv[j] = L[o + j * B];
        sv[B + j] = D[o + j * B];
        sv[2 * B + j] = U[o + j * B];
        L[o + j * B] = 0.0;
        U[o + j * B] = 0.0;
        D[o + j * B] = (j == i) ? 1.0 : 0.0;
    }
    sv[3 * B] = b[h];
    b[h] = rl / cL;
}

// Level s: factor the eliminated blocks i = s + 2s·t and apply D_i⁻¹ to L_i, U_i, b_i.
template <int B>
__global__ __launch_bounds__(kCB) void factor_kernel(int64_t nb, int64_t s, double* __restrict__ D,
                                                     double* __restrict__ L, double* __restrict__ U,
                                                     double* __restrict__ b, unsigned* __restrict__ flag) {
    const int64_t t = (int64_t)blockIdx.x * kCB + threadIdx.x;
    const int64_t i = s + 2 * s * t;
    if (i >= nb) return;
    double a[B][B];
    int piv[B];
    const double* Di = D + i * B * B;
#pragma unroll
    for (int c = 0; c < B; ++c)
#pragma unroll
        for (int r = 0; r < B; ++r) a[r][c] = Di[c * B + r];
    if (!lu_factor<B>(a, piv)) {
        atomicOr(flag, kSingular);
        atomicCAS(status_where(flag), ~0ull, (unsigned long long)i);
    }
    double v[B];
    double* bi = b + i * B;
#pragma unroll
    for (int r = 0; r < B; ++r) v[r] = bi[r];
    lu_solve<B>(a, piv, v);
#pragma unroll
    for (int r = 0; r < B; ++r) bi[r] = v[r];
#pragma unroll 1
    for (int m = 0; m < 2; ++m) {
        double* M = (m == 0 ? L : U) + i * B * B;
#pragma unroll 1
        for (int c = 0; c < B; ++c) {
#pragma unroll
            for (int r = 0; r < B; ++r) v[r] = M[c * B + r];
            lu_solve<B>(a, piv, v);
#pragma unroll
            for (int r = 0; r < B; ++r) M[c * B + r] = v[r];
        }
    }
}

// Level s: one lane per row r of a kept block j = 2s·t.
template <int B>
__global__ __launch_bounds__(kCB) void update_kernel(int64_t nb, int64_t s, double* __restrict__ D,
                                                     double* __restrict__ L, double* __restrict__ U,
                                                     double* __restrict__ b) {
    const int64_t g = (int64_t)blockIdx.x * kCB + threadIdx.x;
    const int64_t t = g / B;
    const int r = (int)(g - t * B);
    const int64_t j = 2 * s * t;
    if (j >= nb) return;
    const int64_t o = j * B * B + r;       // row r of block j: entries o + c·B
    double Lr[B], Ur[B], Dr[B], nL[B], nU[B];
#pragma unroll
    for (int c = 0; c < B; ++c) {
        Lr[c] = L[o + c * B]; Ur[c] = U[o + c * B]; Dr[c] = D[o + c * B]; nL[c] = 0.0; nU[c] = 0.0;
    }
    double br = b[j * B + r];
    if (j >= s) {                          // left neighbour l = j − s (eliminated)
        const int64_t l = j - s;
        const double* Ul = U + l * B * B;
        const double* Ll = L + l * B * B;
        const double* bl = b + l * B;
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const double f = Lr[k];
            br -= f * bl[k];
#pragma unroll
            for (int c = 0; c < B; ++c) {
                Dr[c] -= f * Ul[c * B + k];
                nL[c] -= f * Ll[c * B + k];
            }
        }
    }
    if (j + s < nb) {                      // right neighbour q = j + s (eliminated)
        const int64_t q = j + s;
        const double* Lq = L + q * B * B;
        const double* Uq = U + q * B * B;
        const double* bq = b + q * B;
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const double f = Ur[k];
            br -= f * bq[k];
#pragma unroll
            for (int c = 0; c < B; ++c) {
                Dr[c] -= f * Lq[c * B + k];
                nU[c] -= f * Uq[c * B + k];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < B; ++c) { D[o + c * B] = Dr[c]; L[o + c * B] = nL[c]; U[o + c * B] = nU[c]; }
    b[j * B + r] = br;
}

// The last remaining block: x_0 = D_0⁻¹ b_0.
template <int B>
__global__ void top_kernel(double* __restrict__ D, double* __restrict__ b, unsigned* __restrict__ flag) {
    if (threadIdx.x != 0) return;
    double a[B][B], v[B];
    int piv[B];
#pragma unroll
    for (int r = 0; r < B; ++r) {
        v[r] = b[r];
#pragma unroll
        for (int c = 0; c < B; ++c) a[r][c] = D[c * B + r];
    }
    if (!lu_factor<B>(a, piv)) {
        atomicOr(flag, kSingular);
        atomicCAS(status_where(flag), ~0ull, 0ull);
    }
    lu_solve<B>(a, piv, v);
#pragma unroll
    for (int r = 0; r < B; ++r) b[r] = v[r];
}

// Level s, descending: one lane per row of an eliminated block i = s + 2s·t.
template <int B>
__global__ __launch_bounds__(kCB) void back_kernel(int64_t nb, int64_t s, const double* __restrict__ L,
                                                   const double* __restrict__ U, double* __restrict__ b) {
    const int64_t g = (int64_t)blockIdx.x * kCB + threadIdx.x;
    const int64_t t = g / B;
    const int r = (int)(g - t * B);
    const int64_t i = s + 2 * s * t;
    if (i >= nb) return;
    const int64_t o = i * B * B + r;
    double x = b[i * B + r];
    const double* xl = b + (i - s) * B;
#pragma unroll
    for (int k = 0; k < B; ++k) x -= L[o + k * B] * xl[k];
    if (i + s < nb) {
        const double* xr = b + (i + s) * B;
#pragma unroll
        for (int k = 0; k < B; ++k) x -= U[o + k * B] * xr[k];
    }
    b[i * B + r] = x;
}

// Δx in dof order: x at the dof's position, or λ from the saved row of its held dof.
template <int B>
__global__ __launch_bounds__(kCB) void unpack_kernel(int64_t nX, int64_t nb, const int32_t* __restrict__ pos,
                                                     const int32_t* __restrict__ hpos, const double* __restrict__ pc,
                                                     const double* __restrict__ saved, const double* __restrict__ x,
                                                     double* __restrict__ dx) {
    const int64_t c = (int64_t)blockIdx.x * kCB + threadIdx.x;
    if (c >= nX) return;
    const int32_t pc_ = pos[c];
    if (pc_ >= 0) { dx[c] = x[pc_]; return; }
    const int64_t p = -pc_ - 1;
    const int64_t h = hpos[p], blk = h / B;
    const double* sv = saved + p * (3 * B + 1);
    double r = sv[3 * B];
    if (blk > 0)
        for (int k = 0; k < B; ++k) r -= sv[k] * x[(blk - 1) * B + k];
    for (int k = 0; k < B; ++k) r -= sv[B + k] * x[blk * B + k];
    if (blk + 1 < nb)
        for (int k = 0; k < B; ++k) r -= sv[2 * B + k] * x[(blk + 1) * B + k];
    dx[c] = r / pc[4 * p + 1];
}

// Newmarkβdecrement!{OX} (src/SweepX.jl:98-132) over all X dofs, in the host's operation
// order and without contraction, so that a given Δx decrements the state bitwise as the
// host loop does; per-block partial sums of Δx² and Lλ².
// k: the Newmark coefficients (a₁,a₂,a₃,b₁,b₂,b₃) in device memory, so that a captured graph
// replays with the current ones.
__global__ __launch_bounds__(kCB) void newmark_kernel(int64_t nX, int OX, int first, const double* __restrict__ k,
                                                      const double* __restrict__ dx, const double* __restrict__ rhs,
                                                      const double* __restrict__ scale, double* __restrict__ X0,
                                                      double* __restrict__ X1, double* __restrict__ X2,
                                                      double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double s1[kCB], s2[kCB];
    const double a1 = k[0], a2 = k[1], a3 = k[2], b1 = k[3], b2 = k[4], b3 = k[5];
    const int64_t c = (int64_t)blockIdx.x * kCB + threadIdx.x;
    double q1 = 0.0, q2 = 0.0;
    if (c < nX) {
        const double d = dx[c], sc = scale[c], r = rhs[c];
        q1 = d * d;
        q2 = r * r;
        if (OX == 2) {
            double d1, d2;
            if (first) {
                const double x1 = X1[c] / sc, x2 = X2[c] / sc;
                const double a = a2 * x1 + a3 * x2, bb = b2 * x1 + b3 * x2;
                d1 = a1 * d + a;
                d2 = b1 * d + bb;
            } else {
                d1 = a1 * d;
                d2 = b1 * d;
            }
            X0[c] -= d * sc;
            X1[c] -= d1 * sc;
            X2[c] -= d2 * sc;
        } else if (OX == 1) {
            double d1;
            if (first) {
                const double a = a2 * (X1[c] / sc);
                d1 = a1 * d + a;
            } else {
                d1 = a1 * d;
            }
            X0[c] -= d * sc;
            X1[c] -= d1 * sc;
        } else {
            X0[c] -= d * sc;
        }
    }
    s1[threadIdx.x] = q1;
    s2[threadIdx.x] = q2;
    __syncthreads();
    for (int w = kCB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s1[threadIdx.x] += s1[threadIdx.x + w]; s2[threadIdx.x] += s2[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s1[0]; part[2 * blockIdx.x + 1] = s2[0]; }
}
// One workgroup: the partial sums in a fixed order (deterministic).
__global__ __launch_bounds__(kCB) void norm_kernel(int64_t npart, const double* __restrict__ part,
                                                   double* __restrict__ out) {
    __shared__ double s1[kCB], s2[kCB];
    double q1 = 0.0, q2 = 0.0;
    for (int64_t k = threadIdx.x; k < npart; k += kCB) { q1 += part[2 * k]; q2 += part[2 * k + 1]; }
    s1[threadIdx.x] = q1;
    s2[threadIdx.x] = q2;
    __syncthreads();
    for (int w = kCB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { s1[threadIdx.x] += s1[threadIdx.x + w]; s2[threadIdx.x] += s2[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = s1[0]; out[1] = s2[0]; }
}

}  // namespace chain
