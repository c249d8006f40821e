// MI355X element-evaluation-and-assembly library (implementation of include/muscade_hip.h).
//
// Device data layout
//   node coordinates   xyz[3*n + c]                      (node-major: one 24-byte gather per node)
//   element dof maps   idx[k*nele + e]   int32, 0-based  (SoA: a wave loads 64 consecutive ints)
//   beam data          o2[c*nele + e], sec[f*nele + e]   (SoA)
//   element outputs    static beams (OX = 0): Kbuf[kbase + 45*e + slot], the 45 numbers of
//                      the compact static tangent (beam_math.hpp), element-major; dynamic
//                      beams (OX > 0) and other kinds: Kbuf[kbase + nX²*e + i + nX*j];
//                      Rbuf[rbase + i*nele + e] (SoA), dRbuf likewise (∂R/∂δr of :step)
//   assembly maps      ksrc[p] = Kbuf index of the p-th contribution in (column,row)
//                      order; with compact groups (OX = 0) it is (code << 29) | index,
//                      code = exact multiplier ±1, ±½, ±¼;
//                      kseg[k] = first contribution of nz k; rsrc/rseg likewise for the rhs.
//
// Assembly is a deterministic segmented gather-reduce: nzval[k] = Σ mult·Kbuf[idx],
// p = kseg[k]..kseg[k+1]-1, summed from 0 in increasing p. The radix sort that
// builds ksrc is stable and its input is in Muscade's traversal order
// (ieletyp, iele, jeledof, ieledof — src/Assemble.jl:384-395), so each nz is
// summed in the element order of Muscade's sequential add_∂! loop
// (src/Assemble.jl:470-505,572-586). No atomics touch any assembled value.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <algorithm>
#include <climits>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/muscade_hip.h"
#include "beam_math.hpp"
#include "beam_dyn.hpp"
#include "bar_math.hpp"
#include "chain_solve.hpp"

namespace {

struct Group {
    int kind = 0;
    int64_t nele = 0;
    int nX = 0, nU = 0;
    double scaleX[12];
    double scaleL[12];           // scale.Λ (AssemblyDirect's Lagrangian path)
    double scaleU[3];
    int32_t* d_idx = nullptr;    // nX × nele
    int32_t* d_uidx = nullptr;   // nU × nele
    int32_t* d_nodes = nullptr;  // 2 × nele
    double* d_o2 = nullptr;      // 3 × nele
    double* d_sec = nullptr;     // 16 × nele (beams), 10 × nele (bars)
    double* d_vals = nullptr;    // nele (DofLoad values)
    std::vector<int64_t> h_idx;  // nX × nele, 1-based (for mh_get_asm)
    int64_t kbase = 0, rbase = 0;
    int64_t kper = 0;            // Kbuf entries per element (45 for static beams, nX² otherwise)
    int64_t kcbase = 0;          // dynamic beams: offset of the compact static part in Kc
    int64_t ubase = 0;           // U-dof groups: offset of the per-element ∂L/∂U entries
    int64_t kdbase = 0, kubase = 0;   // AssemblyDirect: offsets of the ∂R/∂X⁽ᵈ⁾ and ∂R/∂U blocks
    bool compact = false;
    // EulerBeam3DA: the stiffnesses of mask bits k (EA, EI₂, EI₃, GJ) are scaled by 10^A[aidx[e]]
    int amask = 0;
    int32_t* d_aidx = nullptr;   // 1 × nele, 0-based A-dof of each element
    double* d_sec_base = nullptr;   // 16 × nele: the section before the scaling (d_sec holds the scaled one)
    int64_t abase = 0;           // offset of the per-element ∂L/∂A entries
};

// One CSC pattern with its contribution maps (seg: first contribution of each nz; src:
// element-block index of each contribution; srcT: the same for the transposed block).
struct DMap {
    int64_t nrow = 0, ncol = 0, nnz = 0, npair = 0;
    int32_t* colptr = nullptr;
    int32_t* rowval = nullptr;
    int32_t* seg = nullptr;
    uint32_t* src = nullptr;
    uint32_t* srcT = nullptr;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t nX = 0, nU = 0, nA = 0, nnod = 0;
    double* d_xyz = nullptr;
    std::vector<Group> groups;
    double* d_X[3] = {nullptr, nullptr, nullptr};
    double* d_U = nullptr;
    double* d_A = nullptr;
    double t = 0.;               // state.time
    // prepared
    bool prepared = false;
    int OX = 0;                  // order of the AssemblySweepX{OX} the maps were prepared for
    bool coded = false;          // ksrc carries multiplier codes (compact groups present)
    int64_t nK = 0, nP = 0, nR = 0, nnz = 0;
    double* d_Kc = nullptr;      // dynamic beams: scaled compact static part, 45 per element
    double* d_Kb = nullptr;      // dynamic beams (SweepX): ε-free kinematics, mh::kDynBase per element
    double* d_ybuf = nullptr;    // dynamic beams (adjoint): y = (J_x Λ, J_v Λ), 15 per element
    double* d_c6 = nullptr;      // SweepX: the Newmark coefficients of the current call (6)
    double* h_c6 = nullptr;      // pinned staging of d_c6
    double* d_Kbuf = nullptr;
    double* d_Rbuf = nullptr;
    double* d_dRbuf = nullptr;
    double* d_nzval = nullptr;
    double* d_rhs = nullptr;
    int32_t* d_colptr = nullptr;
    int32_t* d_rowval = nullptr;
    int32_t* d_kseg = nullptr;
    uint32_t* d_ksrc = nullptr;
    int32_t* d_rseg = nullptr;
    uint32_t* d_rsrc = nullptr;
    unsigned long long* d_flag = nullptr;
    unsigned long long* h_flag = nullptr;   // pinned host copy of d_flag (copied on the stream)
    void* d_scratch = nullptr;   // this device's address of the fixed `scratch` array (set at create)
    // adjoint (VJP) of the residual Lagrangian
    int64_t nRU = 0;             // per-element ∂L/∂U entries
    int32_t* d_useg = nullptr;
    uint32_t* d_usrc = nullptr;
    double* d_lam = nullptr;     // Λ (nX)
    double* d_gbuf = nullptr;    // per-element ∂L/∂X⁽ᵈ⁾ entries, 3 × nR (Rbuf layout per order)
    double* d_gX = nullptr;      // 3 × nX
    double* d_gubuf = nullptr;   // nRU
    double* d_gU = nullptr;      // nU
    int64_t nRA = 0;             // per-element ∂L/∂A entries (EulerBeam3DA elements)
    int32_t* d_aseg = nullptr;   // nA + 1: first entry of each A-dof in the sorted A map
    uint32_t* d_asrc = nullptr;  // nRA: per-element entry of each contribution, by A-dof then element
    double* d_gabuf = nullptr;   // nRA
    double* d_gA = nullptr;      // nA
    double* d_crhs = nullptr;    // nX: right-hand side of an adjoint chain solve
    int64_t* d_spi = nullptr;    // sparse right-hand side: 0-based dofs and values (sp_cap each)
    double* d_spv = nullptr;
    int64_t sp_cap = 0;
    int vjp_nd = 0;              // derivative orders of the last VJP
    // AssemblyDirect{OX,OU,IA} (DirectXUA, one step)
    bool direct = false;
    int dOX = 0, dOU = 0, dIA = 0;
    DMap dXX, dXU, dUX, dUU;     // patterns of (Λ|X, Λ|X), (Λ|X, U), (U, Λ|X), (U, U)
    int64_t nKd = 0, nKU = 0;
    double* d_Kd = nullptr;      // (OX+1) × nKd: per-element ∂R/∂X⁽ᵈ⁾ blocks (× scale of the column, × Δt)
    double* d_KU = nullptr;      // nKU: per-element ∂R/∂U blocks
    double* d_Rd = nullptr;      // nR: per-element L1[Λ] entries
    double* d_L1x = nullptr;     // nR: per-element L1[X][1] entries (Lagrangian-path elements)
    double* d_dL1 = nullptr;     // 2·nX: L1[Λ][1], L1[X][1]
    double* d_dL2 = nullptr;     // 2(OX+1)·nnz(XX) + 2·nnz(XU): L2[Λ,X][1,d], L2[X,Λ][d,1], L2[Λ,U][1,1], L2[U,Λ][1,1]
    int scatter_variant = 0;     // nz scatter shape picked at prepare (launch_scatter)
    int beam_variant = 1;        // static beam kernel shape picked by tune_beams (launch_beam_static)
    float beam_ms[3] = {0.f, 0.f, 0.f};
    bool beam_tuned = false;
    float scatter_ms[16] = {};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool timed = false;
    bool graph_timed = false;    // the last timings are of a graph launch (total only)
    // captured SweepX assembly (mh_assemble_sweepx_graph): replayed while its launch
    // arguments (OX, mission, Newmark coefficients, time, dof scales) are unchanged
    // (one slot per mission: :step and :iter alternate within a time step)
    // captured graphs: slots 0/1 the assembly (:iter/:step), 2/3 a whole device Newton iteration
    // (later/first iteration of a step)
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
    double gkey[4][9] = {};
    int gOX[4] = {-1, -1, -1, -1};
    uint64_t gen = 0, ggen[4] = {~0ull, ~0ull, ~0ull, ~0ull};   // gen: bumped by setters that change launch arguments
    int64_t graph_captures = 0;
    // on-device solve of Lλx·Δx = Lλ for chain topologies (mh_chain_prepare), and the
    // Newton-Newmark decrement (mh_newton_update)
    int cB = 0;                  // block size (3 or 6); 0 = not prepared
    int64_t cnb = 0, cnp = 0, cnd = 0;   // blocks, λ pairs, padding positions
    int32_t* d_cpos = nullptr;   // nX: block position of each dof, or −(p+1) for the λ of pair p
    int32_t* d_chpos = nullptr;  // cnp: block position of each pair's held dof
    int32_t* d_cdum = nullptr;   // cnd: padding positions
    int32_t* d_cslot = nullptr;  // nnz: offset of each nz in the chain buffer (D|L|U, or pc for a λ coupling)
    int32_t* d_cunc = nullptr;   // cnu: the slots no nz covers (zeroed by each solve)
    int64_t cnu = 0;
    double* d_cbuf = nullptr;    // D | L | U (3·cnb·B²) | b (cnb·B) | pair couplings (4·cnp)
    double* d_csaved = nullptr;  // cnp × (3B+1): original rows of the held dofs
    double* d_dx = nullptr;      // nX: Δx of the last solve
    unsigned* d_cflag = nullptr;
    unsigned* h_cflag = nullptr; // pinned
    double* d_scaleX = nullptr;  // nX: dof scales (mh_newton_prepare)
    double* d_part = nullptr;    // 2 × workgroups of newmark_kernel
    double* d_norm = nullptr;
    double* h_norm = nullptr;    // pinned: Σ Δx², Σ Lλ²
    hipEvent_t cev[4] = {nullptr, nullptr, nullptr, nullptr};
    // DirectXUA's all-steps system assembled on the device (mh_big_prepare / mh_big_assemble)
    int64_t bnv = 0, bnp = 0, bnnz = 0, bnhost = 0, bndev = 0;
    int32_t* d_bcolptr = nullptr;
    int32_t* d_browval = nullptr;
    int32_t* d_bseg = nullptr;
    int64_t* d_bsrc = nullptr;   // per contribution, sorted: index into d_dL2 (< bndev) or the host values
    double* d_bw = nullptr;      // per contribution, sorted: its weight (finite-difference factors)
    double* d_bhost = nullptr;   // bnhost host-provided values
    double* d_bval = nullptr;    // bnnz assembled values
    // its dense block solve (mh_big_solve_prepare / mh_big_solve): g steps per diagonal block
    // of order m, nb blocks, na border unknowns (the A-dofs)
    int64_t sps = 0, snst = 0, sna = 0, sg = 0, sm = 0, snb = 0, stot = 0, snpad = 0;
    int64_t so[4] = {0, 0, 0, 0};   // offsets of the band, W, C and E in d_sbuf
    int64_t* d_sslot = nullptr;  // per nz of the big pattern: its position in d_sbuf
    int32_t* d_srowptr = nullptr;  // the pattern by rows (residual of the refinement steps)
    int32_t* d_scolidx = nullptr;
    int32_t* d_sperm = nullptr;    // row-ordered nz → its index in the column-ordered values
    double* d_sb = nullptr;        // the right-hand side, kept
    double* d_sx = nullptr;        // the solution being refined
    int32_t* d_scol = nullptr;     // column of each nz (column-ordered)
    double* d_ssr = nullptr;       // row and column equilibration scales
    double* d_ssc = nullptr;
    int64_t* d_spad = nullptr;   // padding diagonal positions of the last block
    double* d_sbuf = nullptr;
    double* d_srhs = nullptr;
    int32_t* d_sipiv = nullptr;
    int32_t* d_sinfo = nullptr;
    int32_t* d_sprm = nullptr;     // per block column: the row permutation of its interchanges
    rocblas_handle blas = nullptr;
};

#define HIPCHK(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                    \
            return MH_ERR_HIP;                                                                  \
        }                                                                                       \
    } while (0)

template <class T> void dfree(T*& p) { if (p) { (void)hipFree(p); p = nullptr; } }

void free_graph(Ctx* c, int slot) {
    if (c->gexec[slot]) { (void)hipGraphExecDestroy(c->gexec[slot]); c->gexec[slot] = nullptr; }
    if (c->graph[slot]) { (void)hipGraphDestroy(c->graph[slot]); c->graph[slot] = nullptr; }
    c->gOX[slot] = -1;
}
void free_graph(Ctx* c) { for (int k = 0; k < 4; ++k) free_graph(c, k); }
void free_chain(Ctx* c) {
    dfree(c->d_cpos); dfree(c->d_chpos); dfree(c->d_cdum); dfree(c->d_cbuf); dfree(c->d_csaved); dfree(c->d_dx);
    dfree(c->d_cslot); dfree(c->d_cunc); c->cnu = 0;
    dfree(c->d_cflag);
    if (c->h_cflag) { (void)hipHostFree(c->h_cflag); c->h_cflag = nullptr; }
    for (auto& e : c->cev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    c->cB = 0; c->cnb = c->cnp = c->cnd = 0;
}
void free_bsolve(Ctx* c) {
    dfree(c->d_sslot); dfree(c->d_spad); dfree(c->d_sbuf); dfree(c->d_srhs); dfree(c->d_sipiv); dfree(c->d_sinfo);
    dfree(c->d_srowptr); dfree(c->d_scolidx); dfree(c->d_sperm); dfree(c->d_sb); dfree(c->d_sx);
    dfree(c->d_scol); dfree(c->d_ssr); dfree(c->d_ssc); dfree(c->d_sprm);
    c->sps = c->snst = c->sna = c->sg = c->sm = c->snb = c->stot = c->snpad = 0;
}
void free_big(Ctx* c) {
    free_bsolve(c);
    dfree(c->d_bcolptr); dfree(c->d_browval); dfree(c->d_bseg); dfree(c->d_bsrc); dfree(c->d_bw); dfree(c->d_bhost);
    dfree(c->d_bval);
    c->bnv = c->bnp = c->bnnz = c->bnhost = c->bndev = 0;
}
void free_prepared(Ctx* c) {
    free_graph(c);
    free_chain(c);
    free_big(c);
    dfree(c->d_Kbuf); dfree(c->d_Rbuf); dfree(c->d_dRbuf); dfree(c->d_nzval); dfree(c->d_rhs); dfree(c->d_Kc); dfree(c->d_Kb); dfree(c->d_ybuf);
    dfree(c->d_colptr); dfree(c->d_rowval); dfree(c->d_kseg); dfree(c->d_ksrc); dfree(c->d_rseg); dfree(c->d_rsrc);
    dfree(c->d_useg); dfree(c->d_usrc); dfree(c->d_gbuf); dfree(c->d_gX); dfree(c->d_gubuf); dfree(c->d_gU);
    dfree(c->d_aseg); dfree(c->d_asrc); dfree(c->d_gabuf); dfree(c->d_gA); dfree(c->d_crhs);
    dfree(c->d_spi); dfree(c->d_spv); c->sp_cap = 0;
    c->nRA = 0;
    c->vjp_nd = 0;
    for (DMap* M : {&c->dXX, &c->dXU, &c->dUX, &c->dUU}) {
        dfree(M->colptr); dfree(M->rowval); dfree(M->seg); dfree(M->src); dfree(M->srcT);
        *M = DMap();
    }
    dfree(c->d_Kd); dfree(c->d_KU); dfree(c->d_Rd); dfree(c->d_L1x); dfree(c->d_dL1); dfree(c->d_dL2);
    c->direct = false;
    c->prepared = false;
}
void free_model(Ctx* c) {
    free_prepared(c);
    for (auto& g : c->groups) {
        dfree(g.d_idx); dfree(g.d_uidx); dfree(g.d_nodes); dfree(g.d_o2); dfree(g.d_sec); dfree(g.d_vals);
        dfree(g.d_aidx); dfree(g.d_sec_base);
    }
    c->groups.clear();
    dfree(c->d_xyz);
    for (auto& x : c->d_X) dfree(x);
    dfree(c->d_U); dfree(c->d_A); dfree(c->d_lam);
    dfree(c->d_scaleX); dfree(c->d_part); dfree(c->d_norm);
    if (c->h_norm) { (void)hipHostFree(c->h_norm); c->h_norm = nullptr; }
}

// ------------------------------------------------------------------ kernels
constexpr int kBlock = 256;

constexpr int kBeamBlock = 128;
constexpr int kXCD = 8;   // XCDs of an MI355X; workgroups are dispatched round-robin over them

// Workgroup id → logical tile such that consecutive tiles run on one XCD (its own L2).
__device__ __forceinline__ int64_t xcd_tile(int64_t per_xcd) {
    return (int64_t)(blockIdx.x % kXCD) * per_xcd + blockIdx.x / kXCD;
}

struct BeamArgs {
    int64_t nele, per_xcd;
    const int32_t* idx;
    const int32_t* uidx;
    const int32_t* nodes;
    const double* xyz;
    const double* o2;
    const double* sec;
    const double* X;
    const double* U;
    double sA[6], sT[18], sS[21], sR[12];   // dof-scale products per compact slot / residual entry
    double* K;   // compact blocks, 45 per element
    double* R;
    unsigned long long* flag;
    unsigned long long gtag;
};

// One element per lane. Static EulerBeam3D residual + compact tangent (OX=0).
// The 45 tangent numbers of the block's elements are staged in LDS and written
// as one contiguous, coalesced range.
template <bool UDOF, int BLK>
__global__ __launch_bounds__(BLK) void beam_static_kernel(BeamArgs a) {
    __shared__ double tile[BLK * mh::kCompact];
    const int64_t e0 = xcd_tile(a.per_xcd) * BLK;
    if (e0 >= a.nele) return;   // whole workgroup leaves together
    const int tid = threadIdx.x;
    const int64_t e = e0 + tid;
    const int64_t n = a.nele;
    if (e < n) {
        // per-element streams are read once (nontemporal: they leave the caches to the node
        // data, gathered again by neighbouring elements, and to the compact tangents written
        // here for the scatter); X and coordinates are gathers and stay cached
        double X[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) X[k] = a.X[__builtin_nontemporal_load(a.idx + k * n + e)];
        const int32_t n1 = __builtin_nontemporal_load(a.nodes + e), n2 = __builtin_nontemporal_load(a.nodes + n + e);
        mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
        mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
        mh::v3 o2 = mh::mk(__builtin_nontemporal_load(a.o2 + e), __builtin_nontemporal_load(a.o2 + n + e),
                           __builtin_nontemporal_load(a.o2 + 2 * n + e));
        mh::BeamSection S;
        S.EA = __builtin_nontemporal_load(a.sec + e); S.EI2 = __builtin_nontemporal_load(a.sec + n + e);
        S.EI3 = __builtin_nontemporal_load(a.sec + 2 * n + e); S.GJ = __builtin_nontemporal_load(a.sec + 3 * n + e);
        S.w = __builtin_nontemporal_load(a.sec + 6 * n + e);
        mh::v3 fe = mh::mk(0., 0., S.w);
        if (UDOF) {
            fe.x = fe.x - a.U[a.uidx[e]];
            fe.y = fe.y - a.U[a.uidx[n + e]];
            fe.z = fe.z - a.U[a.uidx[2 * n + e]];
        }
        mh::BeamGeo G = mh::beam_geometry(c1, c2, o2);
        mh::BeamOut out;
        mh::beam_static(G, S, fe, X, out);
#pragma unroll
        for (int i = 0; i < 12; ++i) a.R[i * n + e] = out.R[i] * a.sR[i];
        double* t = tile + tid * mh::kCompact;
#pragma unroll
        for (int i = 0; i < 6; ++i) t[i] = out.Kc[i] * a.sA[i];
#pragma unroll
        for (int i = 0; i < 18; ++i) t[6 + i] = out.Kc[6 + i] * a.sT[i];
#pragma unroll
        for (int i = 0; i < 21; ++i) t[24 + i] = out.Kc[24 + i] * a.sS[i];
        if (out.nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    }
    __syncthreads();
    const int64_t nv = (n - e0 < BLK ? n - e0 : BLK) * mh::kCompact;
    double* dst = a.K + e0 * mh::kCompact;
    for (int64_t f = tid; f < nv; f += BLK) dst[f] = tile[f];
}
// The static beam kernel's shapes for the tuner: elements per workgroup, i.e. the LDS tile of
// compact tangents staged for one contiguous write (64: one wave per workgroup).
constexpr int kBeamVariants = 3;
constexpr int kBeamVariantBlock[kBeamVariants] = {64, 128, 256};
void launch_beam_static(int variant, BeamArgs& a, bool udof, hipStream_t s) {
    const int blk = kBeamVariantBlock[variant];
    const int64_t ntile = (a.nele + blk - 1) / blk;
    a.per_xcd = (ntile + kXCD - 1) / kXCD;
    const dim3 gr((unsigned)(a.per_xcd * kXCD));
    switch (variant) {
        case 0: if (udof) hipLaunchKernelGGL((beam_static_kernel<true, 64>), gr, dim3(64), 0, s, a);
                else hipLaunchKernelGGL((beam_static_kernel<false, 64>), gr, dim3(64), 0, s, a); break;
        case 1: if (udof) hipLaunchKernelGGL((beam_static_kernel<true, 128>), gr, dim3(128), 0, s, a);
                else hipLaunchKernelGGL((beam_static_kernel<false, 128>), gr, dim3(128), 0, s, a); break;
        default: if (udof) hipLaunchKernelGGL((beam_static_kernel<true, 256>), gr, dim3(256), 0, s, a);
                 else hipLaunchKernelGGL((beam_static_kernel<false, 256>), gr, dim3(256), 0, s, a); break;
    }
}

struct PointArgs {
    int kind;
    int64_t nele;
    int nX;
    const int32_t* idx;
    const double* X;
    const double* vals;
    double s[12];
    double* K;
    double* R;
    double* dR;
    unsigned long long* flag;
    unsigned long long gtag;
};

// Hold, DofLoad and structural-zero elements: element e of one group.
__device__ __forceinline__ void point_element(const PointArgs& a, int64_t e) {
    const int64_t n = a.nele;
    const int nX = a.nX;
    if (a.kind == MH_KIND_HOLD) {
        // DofConstraint{:X} equal, gap(x,t) = x - g(t) with g the host value (0 for Hold):
        // R = (-λ, -gap), ∂R = [[0,-1],[-1,0]]
        double x = a.X[a.idx[e]], lam = a.X[a.idx[n + e]];
        double R0 = -lam, R1 = -(x - a.vals[e]);
        a.R[e] = R0 * a.s[0];
        a.R[n + e] = R1 * a.s[1];
        double* K = a.K + 4 * e;
        K[0] = 0.;
        K[1] = -1. * a.s[0] * a.s[1];   // (i=1,j=0): ∂R₁/∂x
        K[2] = -1. * a.s[1] * a.s[0];   // (i=0,j=1): ∂R₀/∂λ
        K[3] = 0.;
        if (x != x || lam != lam) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    } else if (a.kind == MH_KIND_DOFLOAD) {
        double F = a.vals[e];
        a.R[e] = -F * a.s[0];
        a.K[e] = 0.;   // 1×1 block
        if (F != F) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    } else if (a.kind != MH_KIND_HOST) {   // host groups: R and K come from mh_set_host_blocks
        for (int i = 0; i < nX; ++i) a.R[i * n + e] = 0.;
        for (int i = 0; i < nX * nX; ++i) a.K[(int64_t)nX * nX * e + i] = 0.;
    }
    if (a.dR && a.kind != MH_KIND_HOST)   // host groups: ∂R/∂δr comes from mh_set_host_blocks / mh_set_host_dr
        for (int i = 0; i < nX; ++i) a.dR[i * n + e] = 0.;
}
// Up to kPointBatch point groups in one launch (they are small; one launch each would
// cost more than their work). Thread q runs element q - start[g] of group g.
constexpr int kPointBatch = 12;
struct PointBatch {
    int ngroup;
    int64_t start[kPointBatch + 1];
    PointArgs g[kPointBatch];
};
__global__ __launch_bounds__(kBlock) void point_kernel(PointBatch b) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.start[b.ngroup]) return;
    int gi = 0;
    while (q >= b.start[gi + 1]) ++gi;
    point_element(b.g[gi], q - b.start[gi]);
}

// ---- dynamic beams (OX = 1, 2) ------------------------------------------------
struct BeamDynArgs {
    int64_t nele, per_xcd;
    const int32_t* idx;
    const int32_t* uidx;
    const int32_t* nodes;
    const double* xyz;
    const double* o2;
    const double* sec;
    const double* X[3];
    const double* U;
    double scale[12];
    double sA[6], sT[18], sS[21], sR[12];
    double c[6];        // Newmark (a₁,a₂,a₃,b₁,b₂,b₃)
    const double* dc = nullptr;   // the same in device memory (SweepX: graphs replay with the current ones)
    double* Kc;         // value pass → direction pass: scaled compact static part, 45 per element
    double* K;          // full 12×12 per element
    double* R;
    double* dR;         // ∂R/∂δr (:step)
    unsigned long long* flag;
    unsigned long long gtag;
    double* Kb = nullptr;   // value pass → direction pass: ε-free kinematics, mh::kDynBase × nele (SoA)
};

template <int ND, bool UDOF>
__device__ __forceinline__ void dyn_load(const BeamDynArgs& a, int64_t e, mh::BeamGeo& G, mh::BeamSection& S,
                                         mh::BeamDynSection& D, double (*X)[12], double* U) {
    const int64_t n = a.nele;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const int32_t q = a.idx[k * n + e];
#pragma unroll
        for (int d = 0; d < ND; ++d) X[d][k] = a.X[d][q];
    }
    const int32_t n1 = a.nodes[e], n2 = a.nodes[n + e];
    mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
    mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
    G = mh::beam_geometry(c1, c2, mh::mk(a.o2[e], a.o2[n + e], a.o2[2 * n + e]));
    S.EA = a.sec[e]; S.EI2 = a.sec[n + e]; S.EI3 = a.sec[2 * n + e]; S.GJ = a.sec[3 * n + e]; S.w = a.sec[6 * n + e];
    D.mu = a.sec[4 * n + e]; D.iota1 = a.sec[5 * n + e]; D.w = S.w;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        D.Ca[q] = a.sec[(7 + 3 * q) * n + e]; D.Cl[q] = a.sec[(8 + 3 * q) * n + e]; D.Cq[q] = a.sec[(9 + 3 * q) * n + e];
        U[q] = UDOF ? a.U[a.uidx[q * n + e]] : 0.;
    }
}

// Value pass, one element per lane: the frozen loads of the time jets, then the
// residual and the compact tangent of Π with those loads held fixed.
template <int ND, bool UDOF>
__global__ __launch_bounds__(kBeamBlock) void beam_dyn_value_kernel(BeamDynArgs a) {
    __shared__ double tile[kBeamBlock * mh::kCompact];
    __shared__ double hmsh[6 * kBeamBlock];   // each lane's couple Hessian, parked (stride kBeamBlock)
    const int64_t e0 = xcd_tile(a.per_xcd) * kBeamBlock;
    if (e0 >= a.nele) return;   // whole workgroup leaves together
    const int tid = threadIdx.x;
    const int64_t e = e0 + tid;
    const int64_t n = a.nele;
    if (e < n) {
        mh::BeamGeo G;
        mh::BeamSection S;
        mh::BeamDynSection D;
        double X[ND][12], U[3];
        dyn_load<ND, UDOF>(a, e, G, S, D, X, U);
        mh::v3 eg[4], mtot;
        mh::dyn_frozen_loads<ND>(G, D, X, U, eg, mtot, a.Kb ? a.Kb + e : nullptr, n);
        mh::BeamOut out;
        mh::beam_static<true>(G, S, mh::mk(0., 0., 0.), X[0], out, eg, mtot, hmsh + tid, kBeamBlock);
        if (a.R) {
#pragma unroll
            for (int i = 0; i < 12; ++i) a.R[i * n + e] = out.R[i] * a.sR[i];
        }
        double* t = tile + tid * mh::kCompact;
#pragma unroll
        for (int i = 0; i < 6; ++i) t[i] = out.Kc[i] * a.sA[i];
#pragma unroll
        for (int i = 0; i < 18; ++i) t[6 + i] = out.Kc[6 + i] * a.sT[i];
#pragma unroll
        for (int i = 0; i < 21; ++i) t[24 + i] = out.Kc[24 + i] * a.sS[i];
        if (out.nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    }
    __syncthreads();
    const int64_t nv = (n - e0 < kBeamBlock ? n - e0 : kBeamBlock) * mh::kCompact;
    double* dst = a.Kc + e0 * mh::kCompact;
    for (int64_t f = tid; f < nv; f += kBeamBlock) dst[f] = tile[f];
}

// Requestables of the beams of one group at the device state (mh_beam_results): one
// element per lane, written element-major (beam_nres(ND) + ngauge numbers per element);
// with strain gauges (E, K1, K2, K3 each, per element) their strains E·ε + K1·κ₁ + K2·κ₂ + K3·κ₃ follow
// (toolbox/StrainGaugeOnBeamElement.jl:70-76).
template <int ND>
__global__ __launch_bounds__(kBlock) void beam_results_kernel(BeamDynArgs a, int ngauge, const double* __restrict__ gauges,
                                                              double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.nele) return;
    mh::BeamGeo G;
    mh::BeamSection S;
    mh::BeamDynSection D;
    double X[ND][12], U[3];
    dyn_load<ND, false>(a, e, G, S, D, X, U);
    constexpr int NR = mh::beam_nres(ND);
    double* o = out + e * (NR + ngauge);
    mh::dyn_requestables<ND>(G, S, D, X, o);
    const double eps = o[0], k1 = o[ND], k2 = o[ND + 1], k3 = o[ND + 2];
    const double* ge = gauges + e * 4 * ngauge;
    for (int k = 0; k < ngauge; ++k)
        o[NR + k] = ((ge[4 * k] * eps + ge[4 * k + 1] * k1) + ge[4 * k + 2] * k2) + ge[4 * k + 3] * k3;
}

// First and second derivatives of the strain-gauge requestables (εₐₓ, κ₁, κ₂, κ₃) w.r.t.
// the 12 element dofs at X₀ (mh_beam_gauge_derivs): one lane per dof pair i ≤ j (78 per
// element). Per element: values (4), gradients (4×12), Hessians (4×78, upper triangle).
constexpr int kPairs = 78;
constexpr int kGaugeOut = 4 + 4 * 12 + 4 * kPairs;
// sel: element numbers (SelT, counted from kSelBase), or null for all elements in order.
template <class SelT, int kSelBase>
__global__ __launch_bounds__(kBlock) void beam_gauge_derivs_kernel(BeamDynArgs a, const SelT* __restrict__ sel,
                                                                   int64_t nsel, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t s = t / kPairs;
    if (s >= nsel) return;
    const int p = (int)(t - s * kPairs);
    int i = 0, rem = p;
    while (rem >= 12 - i) { rem -= 12 - i; ++i; }
    const int j = i + rem;
    const int64_t e = sel ? (int64_t)sel[s] - kSelBase : s;
    mh::BeamGeo G;
    mh::BeamSection S;
    mh::BeamDynSection D;
    double X[1][12], U[3];
    dyn_load<1, false>(a, e, G, S, D, X, U);
    mh::Jet<2, true> q[4];
    mh::gauge_jets(G, X[0], i, j, q);
    double* o = out + s * kGaugeOut;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (p == 0) o[k] = q[k].v[0];
        if (i == j) o[4 + 12 * k + i] = q[k].v[1];
        o[52 + kPairs * k + p] = q[k].e[1];
    }
}

// Direction pass, one wave per workgroup of kDirElemsW(STEP) elements, after the value pass
// (which leaves the compact static part in Kc and the ε-free kinematics in Kb). One lane per
// rotation dof of an element (δX of dofs 3..5, 9..11) and, on :step, one for δr. A lane first
// evaluates the δX direction of one translation dof (0..2, 6..8) in closed form from Kb
// (mh::dyn_translation_direction: along a translation the rotation chain carries no ε), then
// propagates the time jets with its own ε seed. The lanes of an element exchange the Jacobian
// columns of x_gp and vₛₘ (and the translation load derivatives) through LDS, and each lane
// forms the columns k of its directions, K[:,k] = (compact static part) +
// Σ_gp dL·(J_xᵀ ∂fₑ/∂ε + J_vᵀ ∂mₑ/∂ε), in LDS (the workgroup then writes its elements' K blocks
// as one contiguous range), or ∂R/∂δr.
constexpr int kDirLanes = 16;              // the adjoint and DirectXUA kernels: 16 lanes per element
constexpr int kDirElems = 16;
constexpr int kDirBlock = kDirLanes * kDirElems;
constexpr int kDirWBlock = 64;             // SweepX direction pass: one wave
template <bool STEP> constexpr int kDirPerW = STEP ? 7 : 6;             // lanes per element
template <bool STEP> constexpr int kDirElemsW = kDirWBlock / kDirPerW<STEP>;   // 9 on :step, 10 on :iter
template <int ND, bool UDOF, bool STEP>
__global__ __launch_bounds__(kDirWBlock) void beam_dyn_dir_kernel(BeamDynArgs a) {
    constexpr int per = kDirPerW<STEP>, NE = kDirElemsW<STEP>;
    __shared__ double Jsh[NE][12][15];      // per dof: Jx (4 gp × 3), Jv (3)
    __shared__ double DFt[NE][6][12];       // per translation dof: ∂fₑ/∂ε (4 gp × 3); ∂mₑ/∂ε = 0
    __shared__ double Ksh[NE][144];         // the elements' K blocks, written out as one contiguous range
    const int64_t e0 = xcd_tile(a.per_xcd) * NE;
    if (e0 >= a.nele) return;   // whole workgroup leaves together
    const int el = threadIdx.x / per, j = threadIdx.x - el * per;
    const int dir = j < 3 ? 3 + j : (j < 6 ? 6 + j : 12);        // r of node 1, r of node 2, δr
    const int dir_t = j < 3 ? j : j + 3;                          // t of node 1, t of node 2 (j < 6)
    const int64_t e = e0 + el;
    const int64_t n = a.nele;
    const bool live = el < NE && e < n;
    mh::BeamGeo G;
    mh::DynDir d;
    if (live) {
        mh::BeamSection S;
        mh::BeamDynSection D;
        double X[ND][12], U[3];
        dyn_load<ND, UDOF>(a, e, G, S, D, X, U);
        if (j < 6) {
            mh::DynDir dt;
            mh::dyn_translation_direction<ND>(G, D, a.Kb + e, n, a.scale, a.dc ? a.dc : a.c, dir_t, dt);
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int q = 0; q < 3; ++q) { Jsh[el][dir_t][3 * g + q] = dt.Jx[g][q]; DFt[el][j][3 * g + q] = dt.DF[g][q]; }
#pragma unroll
            for (int q = 0; q < 3; ++q) Jsh[el][dir_t][12 + q] = 0.;
        }
        mh::dyn_direction<ND, true>(G, D, X, U, a.scale, a.dc ? a.dc : a.c, dir, d);
        if (dir < 12) {
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int q = 0; q < 3; ++q) Jsh[el][dir][3 * g + q] = d.Jx[g][q];
#pragma unroll
            for (int q = 0; q < 3; ++q) Jsh[el][dir][12 + q] = d.Jv[q];
        }
    }
    __syncthreads();
    if (live) {
        double col[12], colt[12];
        int nan = 0;
#pragma unroll
        for (int i = 0; i < 12; ++i) {
            const double(*Jx)[3] = (const double(*)[3])(&Jsh[el][i][0]);
            col[i] = mh::dyn_row(G, Jx, &Jsh[el][i][12], d);
            if (dir < 12) {
                double t = 0.;
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    t = t + ((Jx[g][0] * DFt[el][j][3 * g] + Jx[g][1] * DFt[el][j][3 * g + 1]) +
                             Jx[g][2] * DFt[el][j][3 * g + 2]) * G.dL[g];
                double m;
                int slot = mh::compact_slot(i, dir, &m);
                col[i] = m * a.Kc[e * mh::kCompact + slot] + col[i];
                slot = mh::compact_slot(i, dir_t, &m);
                colt[i] = m * a.Kc[e * mh::kCompact + slot] + t;
                nan |= colt[i] != colt[i];
            }
            nan |= col[i] != col[i];
        }
        if (dir < 12) {
#pragma unroll
            for (int i = 0; i < 12; ++i) Ksh[el][12 * dir + i] = col[i];
#pragma unroll
            for (int i = 0; i < 12; ++i) Ksh[el][12 * dir_t + i] = colt[i];
        } else {
#pragma unroll
            for (int i = 0; i < 12; ++i) a.dR[i * n + e] = col[i];
        }
        if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    }
    __syncthreads();
    // the workgroup's elements are consecutive: their K blocks form one contiguous range
    const int64_t nv = (n - e0 < NE ? n - e0 : NE) * 144;
    double* dst = a.K + e0 * 144;
    const double* src = &Ksh[0][0];
    for (int64_t f = threadIdx.x; f < nv; f += kDirWBlock) dst[f] = src[f];
}

// ---- Bar3D -----------------------------------------------------------------
struct BarArgs {
    int64_t nele, per_xcd;
    const int32_t* idx;
    const int32_t* uidx;
    const int32_t* nodes;
    const double* xyz;
    const double* sec;   // 10 × nele: AxisymmetricBarCrossSection fields, then ϵₛ
    const double* X[3];
    const double* U;
    double t;
    double scale[6];
    double c[6];
    const double* dc = nullptr;   // the Newmark coefficients in device memory, if set (else c)
    double* K;           // 36 per element, element-major, column-major inside
    double* R;
    double* dR;
    unsigned long long* flag;
    unsigned long long gtag;
};
constexpr int kBarBlock = 128;
// One bar per lane (any OX): residual, K + a₁C + b₁M and, on :step, ∂R/∂δr. The 36
// tangent numbers of the block are staged in LDS and written as one contiguous range.
template <int ND, bool UDOF, bool STEP>
__global__ __launch_bounds__(kBarBlock) void bar_kernel(BarArgs a) {
    __shared__ double tile[kBarBlock * 36];
    const int64_t e0 = xcd_tile(a.per_xcd) * kBarBlock;
    if (e0 >= a.nele) return;   // whole workgroup leaves together
    const int tid = threadIdx.x;
    const int64_t e = e0 + tid;
    const int64_t n = a.nele;
    if (e < n) {
        double X[ND][6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int32_t q = a.idx[k * n + e];
#pragma unroll
            for (int d = 0; d < ND; ++d) X[d][k] = a.X[d][q];
        }
        const int32_t n1 = a.nodes[e], n2 = a.nodes[n + e];
        mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
        mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
        const mh::BarGeo G = mh::bar_geometry(c1, c2, a.sec[9 * n + e]);
        mh::BarSection S;
        S.EA = a.sec[e]; S.mu = a.sec[n + e]; S.w = a.sec[2 * n + e]; S.Cat = a.sec[3 * n + e]; S.Clt = a.sec[4 * n + e];
        S.Cqt = a.sec[5 * n + e]; S.Can = a.sec[6 * n + e]; S.Cln = a.sec[7 * n + e]; S.Cqn = a.sec[8 * n + e];
        double U[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) U[q] = UDOF ? a.U[a.uidx[q * n + e]] : 0.;
        double R[6], dR[6];
        double* K = tile + tid * 36;
        const int nan = mh::bar_element<ND>(G, S, X, U, a.t, a.scale, a.dc ? a.dc : a.c, R, K, STEP ? dR : nullptr);
#pragma unroll
        for (int i = 0; i < 6; ++i) a.R[i * n + e] = R[i];
        if (STEP) {
#pragma unroll
            for (int i = 0; i < 6; ++i) a.dR[i * n + e] = dR[i];
        }
        if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    }
    __syncthreads();
    const int64_t nv = (n - e0 < kBarBlock ? n - e0 : kBarBlock) * 36;
    double* dst = a.K + e0 * 36;
    for (int64_t f = tid; f < nv; f += kBarBlock) dst[f] = tile[f];
}

// ---- adjoint: Λ·∂R/∂X⁽ᵈ⁾, Λ·∂R/∂U per element ---------------------------------
struct VjpArgs {
    int64_t nele, per_xcd;
    int nX;
    int kind;
    const int32_t* idx;
    const int32_t* uidx;
    const int32_t* nodes;
    const double* xyz;
    const double* o2;
    const double* sec;
    const double* X[3];
    const double* U;
    const double* lam;
    double t;
    const double* Kc;    // dynamic beams: unscaled compact static part (value pass)
    const double* Kb;    // dynamic beams: ε-free kinematics (value pass), mh::kDynBase per element
    double* yb;          // dynamic beams: y = (J_x Λ, J_v Λ), 15 per element (yb[q*nele + e])
    double one[12] = {1., 1., 1., 1., 1., 1., 1., 1., 1., 1., 1., 1.};   // unit dof scales
    double* g;           // g[der*gstride + i*nele + e]
    int64_t gstride;
    double* gu;          // gu[q*nele + e] (U-dof groups)
    unsigned long long* flag;
    unsigned long long gtag;
};

// Static beams, one per lane: Λ·K = K·Λ with the compact symmetric tangent; U-dofs
// through y = ∂x_gp/∂X·Λ (one ε-jet pass).
template <bool UDOF>
__global__ __launch_bounds__(kBeamBlock) void beam_vjp_static_kernel(VjpArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nele) return;
    const int64_t n = a.nele;
    double X[1][12], lam[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) { const int32_t q = a.idx[k * n + e]; X[0][k] = a.X[0][q]; lam[k] = a.lam[q]; }
    const int32_t n1 = a.nodes[e], n2 = a.nodes[n + e];
    mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
    mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
    const mh::BeamGeo G = mh::beam_geometry(c1, c2, mh::mk(a.o2[e], a.o2[n + e], a.o2[2 * n + e]));
    mh::BeamSection S;
    S.EA = a.sec[e]; S.EI2 = a.sec[n + e]; S.EI3 = a.sec[2 * n + e]; S.GJ = a.sec[3 * n + e]; S.w = a.sec[6 * n + e];
    mh::v3 fe = mh::mk(0., 0., S.w);
    int nan = 0;
    if (UDOF) {
        fe = fe - mh::mk(a.U[a.uidx[e]], a.U[a.uidx[n + e]], a.U[a.uidx[2 * n + e]]);
        // ∂L/∂U first, so that nothing but λ stays live through the closed form
        mh::DynY y;
        mh::dyn_y(G, X, lam, y);
        const mh::v3 gu = mh::dyn_vjp_u(G, y);
        a.gu[e] = gu.x; a.gu[n + e] = gu.y; a.gu[2 * n + e] = gu.z;
    }
    mh::BeamOut out;
    mh::beam_static(G, S, fe, X[0], out);
    nan |= out.nan;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        double g = 0.;
        for (int j = 0; j < 12; ++j) g = g + mh::compact_entry(out.Kc, i, j) * lam[j];
        a.g[i * n + e] = g;
        nan |= g != g;
    }
    if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
}

// EulerBeam3DA: the scaled section, sec = base · 10^A[aidx[e]] on the stiffnesses of the mask.
__global__ __launch_bounds__(kBlock) void scale_section_kernel(int64_t n, const int32_t* __restrict__ aidx,
                                                               const double* __restrict__ A, int mask,
                                                               const double* __restrict__ base, double* __restrict__ sec) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double f = exp10(A[aidx[e]]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (mask >> k & 1) sec[k * n + e] = base[k * n + e] * f;
}

// ∂L/∂a of EulerBeam3DA elements, L = Λᵀ R: R depends on a only through the scaled
// stiffnesses, linearly, so ∂R/∂a = ln 10 · R(X; the scaled stiffnesses alone, no loads), the
// internal force of those stiffnesses. One element per lane; the internal force is the
// static closed form at X (for the dynamic beam too: the loads do not depend on stiffness).
// VEC: the vector ∂R/∂a itself, 12 per element (element-major) into ga, instead of Λ·∂R/∂a.
template <bool VEC>
__global__ __launch_bounds__(kBeamBlock) void beam_vjp_a_kernel(VjpArgs a, int mask, double* __restrict__ ga) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nele) return;
    const int64_t n = a.nele;
    double X[12], lam[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const int32_t q = a.idx[k * n + e];
        X[k] = a.X[0][q];
        lam[k] = VEC ? 0. : a.lam[q];
    }
    const int32_t n1 = a.nodes[e], n2 = a.nodes[n + e];
    mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
    mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
    const mh::BeamGeo G = mh::beam_geometry(c1, c2, mh::mk(a.o2[e], a.o2[n + e], a.o2[2 * n + e]));
    mh::BeamSection S;
    S.EA = (mask & 1) ? a.sec[e] : 0.;
    S.EI2 = (mask & 2) ? a.sec[n + e] : 0.;
    S.EI3 = (mask & 4) ? a.sec[2 * n + e] : 0.;
    S.GJ = (mask & 8) ? a.sec[3 * n + e] : 0.;
    S.w = 0.;
    mh::BeamOut out;
    mh::beam_static(G, S, mh::mk(0., 0., 0.), X, out);
    constexpr double kLn10 = 2.302585092994045684;
    if constexpr (VEC) {
        int nan = 0;
#pragma unroll
        for (int i = 0; i < 12; ++i) {
            const double v = kLn10 * out.R[i];
            ga[e * 12 + i] = v;
            nan |= v != v;
        }
        if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    } else {
        double g = 0.;
#pragma unroll
        for (int i = 0; i < 12; ++i) g = g + lam[i] * out.R[i];
        g = g * kLn10;
        ga[e] = g;
        if (g != g) atomicMin(a.flag, a.gtag | (unsigned long long)e);
    }
}

// Dynamic beams, after the value pass with unit scales (the compact static part in Kc, the
// ε-free kinematics in Kb), one element per lane: y = (J_x Λ, J_v Λ), the derivatives of x_gp
// (4 × 3) and vₛₘ (3) along Λ (a first-order jet seeded with Λ), 15 numbers per element into
// yb (yb[q·n + e]); the static part (Kc Λ)ₖ of ∂L/∂Xₖ into g; ∂L/∂U = −Σ_gp dL y_gp with U-dofs.
template <int ND, bool UDOF>
__global__ __launch_bounds__(kBlock) void beam_dyn_y_kernel(VjpArgs a) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t n = a.nele;
    if (e >= n) return;
    mh::BeamGeo G;
    mh::BeamSection S;
    mh::BeamDynSection D;
    double X[ND][12], U[3], lam[12];
    {
        BeamDynArgs b;
        b.nele = a.nele; b.idx = a.idx; b.uidx = a.uidx; b.nodes = a.nodes; b.xyz = a.xyz; b.o2 = a.o2; b.sec = a.sec;
        for (int d = 0; d < 3; ++d) b.X[d] = a.X[d];
        b.U = a.U;
        dyn_load<ND, UDOF>(b, e, G, S, D, X, U);
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) lam[i] = a.lam[a.idx[i * n + e]];
    mh::DynY y;
    mh::dyn_y(G, X, lam, y);
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int q = 0; q < 3; ++q) a.yb[(3 * g + q) * n + e] = y.y[g][q];
#pragma unroll
    for (int q = 0; q < 3; ++q) a.yb[(12 + q) * n + e] = y.yv[q];
    if (UDOF) {
        const mh::v3 gu = mh::dyn_vjp_u(G, y);
        a.gu[e] = gu.x; a.gu[n + e] = gu.y; a.gu[2 * n + e] = gu.z;
    }
    const double* kc = a.Kc + e * mh::kCompact;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        double g = 0.;
#pragma unroll
        for (int i = 0; i < 12; ++i) g = g + mh::compact_entry(kc, k, i) * lam[i];
        a.g[k * n + e] = g;
    }
}

// Then one launch per time-derivative order DER: one wave per workgroup of kDirElemsW<false>
// elements, 6 lanes per element, lane j owning one translation dof (0..2, 6..8) and one
// rotation dof (3..5, 9..11) as in the direction pass. A lane differentiates the loads along
// X⁽ᴰᴱᴿ⁾ₖ of its two dofs — the rotation on time jets, then the translation in closed form — and
// contracts with y: ∂L/∂X⁽ᴰᴱᴿ⁾ₖ = Σ_gp dL (y_gp·∂fₑ/∂ε + y_v·∂mₑ/∂ε), added to the static part
// for DER = 0. (All orders in one launch keep three jet passes' state live at once and spill
// about 2 KB per lane.)
template <int ND, bool UDOF, int DER>
__global__ __launch_bounds__(kDirWBlock) void beam_dyn_vjp_kernel(VjpArgs a) {
    constexpr int per = kDirPerW<false>, NE = kDirElemsW<false>;
    const int64_t e0 = xcd_tile(a.per_xcd) * NE;
    const int el = threadIdx.x / per, j = threadIdx.x - el * per;
    const int64_t e = e0 + el;
    const int64_t n = a.nele;
    if (el >= NE || e >= n) return;
    const int dir_t = j < 3 ? j : j + 3;     // t of node 1, t of node 2
    const int dir_r = j < 3 ? 3 + j : 6 + j; // r of node 1, r of node 2
    mh::BeamGeo G;
    mh::BeamSection S;
    mh::BeamDynSection D;
    double X[ND][12], U[3];
    {
        BeamDynArgs b;
        b.nele = a.nele; b.idx = a.idx; b.uidx = a.uidx; b.nodes = a.nodes; b.xyz = a.xyz; b.o2 = a.o2; b.sec = a.sec;
        for (int d = 0; d < 3; ++d) b.X[d] = a.X[d];
        b.U = a.U;
        dyn_load<ND, UDOF>(b, e, G, S, D, X, U);
    }
    // the rotation direction first, while nothing else is live
    const double c[6] = {DER == 1 ? 1. : 0., 0., 0., DER == 2 ? 1. : 0., 0., 0.};
    constexpr double w0 = DER == 0 ? 1. : 0.;
    mh::DynDir d;
    mh::DynY y;
    mh::dyn_direction<ND, true>(G, D, X, U, a.one, c, dir_r, d, w0);
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int q = 0; q < 3; ++q) y.y[g][q] = a.yb[(3 * g + q) * n + e];
#pragma unroll
    for (int q = 0; q < 3; ++q) y.yv[q] = a.yb[(12 + q) * n + e];
    double gr = mh::dyn_ycontract(G, y, d);
    mh::dyn_translation_direction<ND>(G, D, a.Kb + e, n, a.one, c, dir_t, d, w0);
    double gt = mh::dyn_ycontract(G, y, d);
    if constexpr (DER == 0) {   // the static part, from the y kernel
        gt = a.g[dir_t * n + e] + gt;
        gr = a.g[dir_r * n + e] + gr;
    }
    a.g[DER * a.gstride + dir_t * n + e] = gt;
    a.g[DER * a.gstride + dir_r * n + e] = gr;
    if ((gt != gt) | (gr != gr)) atomicMin(a.flag, a.gtag | (unsigned long long)e);
}

template <int ND, bool UDOF>
__global__ __launch_bounds__(kBarBlock) void bar_vjp_kernel(VjpArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nele) return;
    const int64_t n = a.nele;
    double X[ND][6], lam[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int32_t q = a.idx[k * n + e];
        lam[k] = a.lam[q];
#pragma unroll
        for (int d = 0; d < ND; ++d) X[d][k] = a.X[d][q];
    }
    const int32_t n1 = a.nodes[e], n2 = a.nodes[n + e];
    mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
    mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
    const mh::BarGeo G = mh::bar_geometry(c1, c2, a.sec[9 * n + e]);
    mh::BarSection S;
    S.EA = a.sec[e]; S.mu = a.sec[n + e]; S.w = a.sec[2 * n + e]; S.Cat = a.sec[3 * n + e]; S.Clt = a.sec[4 * n + e];
    S.Cqt = a.sec[5 * n + e]; S.Can = a.sec[6 * n + e]; S.Cln = a.sec[7 * n + e]; S.Cqn = a.sec[8 * n + e];
    double U[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) U[q] = UDOF ? a.U[a.uidx[q * n + e]] : 0.;
    double gX[ND][6], gu[3];
    const int nan = mh::bar_vjp_element<ND>(G, S, X, U, a.t, lam, gX, gu);
#pragma unroll
    for (int der = 0; der < ND; ++der)
#pragma unroll
        for (int i = 0; i < 6; ++i) a.g[der * a.gstride + i * n + e] = gX[der][i];
    if (UDOF)
        for (int q = 0; q < 3; ++q) a.gu[q * n + e] = gu[q];
    if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
}

// Hold: L = Λₓ(-λ) + Λ_λ(-x); DofLoad and structural zeros do not depend on X.
__global__ __launch_bounds__(kBlock) void point_vjp_kernel(VjpArgs a, int nd) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nele) return;
    const int64_t n = a.nele;
    for (int der = 0; der < nd; ++der)
        for (int i = 0; i < a.nX; ++i) a.g[der * a.gstride + i * n + e] = 0.;
    if (a.kind == MH_KIND_HOLD) {
        const double lx = a.lam[a.idx[e]], ll = a.lam[a.idx[n + e]];
        a.g[e] = -ll;
        a.g[n + e] = -lx;
    }
}

// ---- AssemblyDirect: per-element blocks ∂R/∂X⁽ᵈ⁾, ∂R/∂U of one DirectXUA step ------
// Residual path (beams, bars; DirectXUA.jl:85-120): R unscaled, partials × scale of the
// variable (revariate{1}((;X,U),(;X=scale.X,U=scale.U))), all × Δt.
struct DirectArgs {
    int64_t nele, per_xcd;
    int nX;
    int kind;
    const int32_t* idx;
    const int32_t* uidx;
    const int32_t* nodes;
    const double* xyz;
    const double* o2;
    const double* sec;
    const double* X[3];
    const double* U;
    const double* lam;
    const double* vals;   // DofLoad values
    double t, dt;
    double sX[12], sU[3], sL[12];
    const double* Kc;     // beams: unscaled compact static part (value pass)
    double* Kd;           // Kd[d*nKd + kdbase + e*nX² + i + nX·k]
    int64_t nKd;
    double* KU;           // KU[kubase + e*nX*3 + i + nX·c]
    double* R;            // per-element L1[Λ] (Rbuf layout)
    double* L1x;          // per-element L1[X][1]
    unsigned long long* flag;
    unsigned long long gtag;
};

// Beams, 16 lanes per element: the X₀ pass of lane k gives column k of J_x, J_v (LDS) and
// of ∂R/∂X₀; lanes then differentiate the loads along X′ₖ and X″ₖ for the other orders;
// lane 12 forms ∂R/∂U = -Σ dL J_xᵀ.
template <int ND, bool UDOF>
__global__ __launch_bounds__(kDirBlock) void beam_direct_kernel(DirectArgs a) {
    __shared__ double Jsh[kDirElems][12][15];
    const int64_t e0 = xcd_tile(a.per_xcd) * kDirElems;
    if (e0 >= a.nele) return;   // whole workgroup leaves together
    const int el = threadIdx.x / kDirLanes, k = threadIdx.x % kDirLanes;
    const int64_t e = e0 + el;
    const int64_t n = a.nele;
    const bool live = e < n && k <= 12;
    mh::BeamGeo G;
    const double one[12] = {1., 1., 1., 1., 1., 1., 1., 1., 1., 1., 1., 1.};
    mh::DynDir d0;
    // the element's data from memory (again for the higher time orders, rather than held
    // live across the barrier)
    auto load = [&](mh::BeamGeo& Gl, mh::BeamDynSection& Dl, double (*Xl)[12], double* Ul) {
        mh::BeamSection S;
        BeamDynArgs b;
        b.nele = a.nele; b.idx = a.idx; b.uidx = a.uidx; b.nodes = a.nodes; b.xyz = a.xyz; b.o2 = a.o2; b.sec = a.sec;
        for (int d = 0; d < 3; ++d) b.X[d] = a.X[d];
        b.U = a.U;
        dyn_load<ND, UDOF>(b, e, Gl, S, Dl, Xl, Ul);
    };
    if (live) {
        mh::BeamDynSection D;
        double X[ND][12], U[3];
        load(G, D, X, U);
        if (k < 12) {
            const double c0[6] = {0., 0., 0., 0., 0., 0.};
            mh::dyn_direction<ND>(G, D, X, U, one, c0, k, d0, 1.);
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int q = 0; q < 3; ++q) Jsh[el][k][3 * g + q] = d0.Jx[g][q];
#pragma unroll
            for (int q = 0; q < 3; ++q) Jsh[el][k][12 + q] = d0.Jv[q];
        }
    }
    __syncthreads();
    if (!live) return;
    int nan = 0;
    if (k == 12) {
        if (UDOF) {
            double* dst = a.KU + e * 36;
            for (int c = 0; c < 3; ++c)
                for (int i = 0; i < 12; ++i) {
                    double s = 0.;
                    for (int g = 0; g < 4; ++g) s = s + Jsh[el][i][3 * g + c] * G.dL[g];
                    dst[i + 12 * c] = ((-s) * a.sU[c]) * a.dt;
                    nan |= dst[i + 12 * c] != dst[i + 12 * c];
                }
        }
    } else {
        // column k of the block of each time order; d0 (order 0) is consumed first, so that it is
        // not live through the directions of the higher orders
        auto rows = [&](int der, const mh::DynDir& d, const mh::BeamGeo& Gx) {
            double* dst = a.Kd + der * a.nKd + e * 144 + 12 * k;
            for (int i = 0; i < 12; ++i) {
                double v = mh::dyn_row(Gx, (const double(*)[3])(&Jsh[el][i][0]), &Jsh[el][i][12], d);
                if (der == 0) {
                    double m;
                    const int slot = mh::compact_slot(i, k, &m);
                    v = m * a.Kc[e * mh::kCompact + slot] + v;
                }
                dst[i] = (v * a.sX[k]) * a.dt;
                nan |= dst[i] != dst[i];
            }
        };
        rows(0, d0, G);
        for (int der = 1; der < ND; ++der) {
            mh::BeamGeo Gd;
            mh::BeamDynSection D;
            double X[ND][12], U[3];
            load(Gd, D, X, U);
            mh::DynDir d;
            const double c[6] = {der == 1 ? 1. : 0., 0., 0., der == 2 ? 1. : 0., 0., 0.};
            mh::dyn_direction<ND>(Gd, D, X, U, one, c, k, d, 0.);
            rows(der, d, Gd);
        }
    }
    if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
}

template <int ND, bool UDOF>
__global__ __launch_bounds__(kBarBlock) void bar_direct_kernel(DirectArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nele) return;
    const int64_t n = a.nele;
    double X[ND][6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int32_t q = a.idx[k * n + e];
#pragma unroll
        for (int d = 0; d < ND; ++d) X[d][k] = a.X[d][q];
    }
    const int32_t n1 = a.nodes[e], n2 = a.nodes[n + e];
    mh::v3 c1 = mh::mk(a.xyz[3 * (int64_t)n1], a.xyz[3 * (int64_t)n1 + 1], a.xyz[3 * (int64_t)n1 + 2]);
    mh::v3 c2 = mh::mk(a.xyz[3 * (int64_t)n2], a.xyz[3 * (int64_t)n2 + 1], a.xyz[3 * (int64_t)n2 + 2]);
    const mh::BarGeo G = mh::bar_geometry(c1, c2, a.sec[9 * n + e]);
    mh::BarSection S;
    S.EA = a.sec[e]; S.mu = a.sec[n + e]; S.w = a.sec[2 * n + e]; S.Cat = a.sec[3 * n + e]; S.Clt = a.sec[4 * n + e];
    S.Cqt = a.sec[5 * n + e]; S.Can = a.sec[6 * n + e]; S.Cln = a.sec[7 * n + e]; S.Cqn = a.sec[8 * n + e];
    double U[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) U[q] = UDOF ? a.U[a.uidx[q * n + e]] : 0.;
    const double one[6] = {1., 1., 1., 1., 1., 1.};
    double R[6], K[36];
    int nan = 0;
    for (int der = 0; der < ND; ++der) {
        const double c[6] = {der == 1 ? 1. : 0., 0., 0., der == 2 ? 1. : 0., 0., 0.};
        nan |= mh::bar_element<ND>(G, S, X, U, a.t, one, c, R, K, nullptr, der == 0 ? 1. : 0.);
        double* dst = a.Kd + der * a.nKd + e * 36;
        for (int k = 0; k < 6; ++k)
            for (int i = 0; i < 6; ++i) dst[i + 6 * k] = (K[i + 6 * k] * a.sX[k]) * a.dt;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) a.R[i * n + e] = R[i] * a.dt;
    if (UDOF) {
        // ∂Rⱼ/∂U_q = -Σ_gp w ψⱼ [j mod 3 = q]
        double* dst = a.KU + e * 18;
        for (int q = 0; q < 3; ++q)
            for (int j = 0; j < 6; ++j) {
                double s = 0.;
                if (j % 3 == q)
                    for (int g = 0; g < 4; ++g) s = s + (j < 3 ? (-G.zgp[g] + 1. / 2.) : (G.zgp[g] + 1. / 2.)) * G.wgp[g];
                dst[j + 6 * q] = ((-s) * a.sU[q]) * a.dt;
            }
    }
    if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
}

// Lagrangian path of Hold and DofLoad (L = Λ∘₁R, revariate{2} with scale.Λ, scale.X;
// DirectXUA.jl:121-171): L1[Λ] = R·sΛ, L1[X] = (Λᵀ∂R/∂X)·sX, L2[Λ,X] = ∂R/∂X·sΛ·sX.
__global__ __launch_bounds__(kBlock) void point_direct_kernel(DirectArgs a, int nd) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nele) return;
    const int64_t n = a.nele;
    const int nX = a.nX, per = nX * nX;
    for (int der = 0; der < nd; ++der)
        for (int q = 0; q < per; ++q) a.Kd[der * a.nKd + e * per + q] = 0.;
    for (int i = 0; i < nX; ++i) { a.R[i * n + e] = 0.; a.L1x[i * n + e] = 0.; }
    int nan = 0;
    if (a.kind == MH_KIND_HOLD) {
        const double x = a.X[0][a.idx[e]], lam = a.X[0][a.idx[n + e]];
        const double Lx = a.lam[a.idx[e]], Ll = a.lam[a.idx[n + e]];
        a.R[e] = (-lam * a.sL[0]) * a.dt;
        a.R[n + e] = (-x * a.sL[1]) * a.dt;
        a.L1x[e] = (-Ll * a.sX[0]) * a.dt;
        a.L1x[n + e] = (-Lx * a.sX[1]) * a.dt;
        a.Kd[e * 4 + 1] = ((-1. * a.sL[1]) * a.sX[0]) * a.dt;   // (i=1, k=0): ∂R₁/∂x
        a.Kd[e * 4 + 2] = ((-1. * a.sL[0]) * a.sX[1]) * a.dt;   // (i=0, k=1): ∂R₀/∂λ
        nan = (x != x) || (lam != lam);
    } else if (a.kind == MH_KIND_DOFLOAD) {
        a.R[e] = (-a.vals[e] * a.sL[0]) * a.dt;
        nan = a.vals[e] != a.vals[e];
    }
    if (nan) atomicMin(a.flag, a.gtag | (unsigned long long)e);
}

// Contribution source: (code << 29) | index; the code selects an exact multiplier.
constexpr uint32_t kIdxMask = (1u << 29) - 1;
__device__ __forceinline__ double contrib(const double* __restrict__ buf, uint32_t s) {
    const double v = buf[s & kIdxMask];
    switch (s >> 29) {
    case 0: return v;
    case 1: return -v;
    case 2: return 0.5 * v;
    case 3: return -0.5 * v;
    case 4: return 0.25 * v;
    default: return -0.25 * v;
    }
}
// Segmented gather-reduce: out[k] = Σ contrib(src[p]), p = seg[k]..seg[k+1]-1, summed
// from zero in increasing p — the order of Muscade's sequential add_∂!/add_value! loop.
// A workgroup owns kTile consecutive segments, whose contributions are contiguous in p:
// the whole workgroup loads them (coalesced indices, every lane gathering) into LDS in
// chunks, then each lane sums its own segment sequentially from LDS. Lanes of long and
// short segments therefore share the memory work evenly.
// STEP (rhs of the :step mission): each contribution adds R then subtracts ∂R/∂δr
// (add_value! then add_∂!{1,:minus}, src/SweepX.jl:56-57).
constexpr int kTile = 256;
constexpr int kChunk = 1024;
// CODED: src carries multiplier codes (compact static beams); otherwise src is the
// plain 32-bit index. SPL segments per lane (lane t owns k0 + t + j·kTile, so the
// writes stay coalesced) and CH contributions per LDS chunk are the shape the
// prepare-time tuner picks (scatter_variant).
// IDX: segment k is written to out[dst[k]] instead of out[k] (the nz straight into the
// chain solve's blocks).
template <bool STEP, bool CODED, int SPL = 1, int CH = kChunk, bool NTM = false, bool IDX = false>
__global__ __launch_bounds__(kTile) void scatter_kernel(int64_t nseg, int64_t per_xcd, const int32_t* __restrict__ seg,
                                                        const uint32_t* __restrict__ src, const double* __restrict__ buf,
                                                        const double* __restrict__ dbuf, double* __restrict__ out,
                                                        const int32_t* __restrict__ dst = nullptr) {
    __shared__ double vals[CH];
    __shared__ double dvals[STEP ? CH : 1];
    constexpr int kSeg = kTile * SPL;
    const int64_t k0 = xcd_tile(per_xcd) * kSeg;
    if (k0 >= nseg) return;   // the whole workgroup leaves together
    const int t = threadIdx.x;
    const int64_t klast = (k0 + kSeg < nseg ? k0 + kSeg : nseg);
    const int32_t pbeg = seg[k0], pend = seg[klast];
    int32_t p0[SPL], p1[SPL];
    double acc[SPL];
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        const int64_t k = k0 + t + j * kTile;
        p0[j] = p1[j] = 0;
        if (k < nseg) {
            if (NTM) { p0[j] = __builtin_nontemporal_load(seg + k); p1[j] = __builtin_nontemporal_load(seg + k + 1); }
            else { p0[j] = seg[k]; p1[j] = seg[k + 1]; }
        }
        acc[j] = 0.;
    }
    for (int32_t c = pbeg; c < pend; c += CH) {
        const int32_t cend = (c + CH < pend ? c + CH : pend);
        for (int32_t q = c + t; q < cend; q += kTile) {
            const uint32_t sq = NTM ? __builtin_nontemporal_load(src + q) : src[q];
            vals[q - c] = CODED ? contrib(buf, sq) : buf[sq];
            if (STEP) dvals[q - c] = dbuf[CODED ? (sq & kIdxMask) : sq];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < SPL; ++j) {
            const int32_t a = p0[j] > c ? p0[j] : c, b = p1[j] < cend ? p1[j] : cend;
            for (int32_t q = a; q < b; ++q) {
                acc[j] += vals[q - c];
                if (STEP) acc[j] -= dvals[q - c];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < SPL; ++j) {
        const int64_t k = k0 + t + j * kTile;
        if (k < nseg) {
            const int64_t o = IDX ? (int64_t)(NTM ? __builtin_nontemporal_load(dst + k) : dst[k]) : k;
            if (NTM) __builtin_nontemporal_store(acc[j], out + o);
            else out[o] = acc[j];
        }
    }
}
// Launch one of the tuned shapes of the nz scatter: (segments per lane, LDS chunk).
// Variants 8..15 are 0..7 with nontemporal index loads and nzval stores (streamed data kept
// out of the caches, which then hold more of the gathered element outputs).
constexpr int kScatterVariants = 16;
constexpr int kScatterSPL[kScatterVariants] = {1, 2, 2, 3, 4, 4, 6, 8, 1, 2, 2, 3, 4, 4, 6, 8};
template <bool CODED, bool IDX>
void launch_scatter_shape(int variant, hipStream_t s, int64_t nseg, const int32_t* seg, const uint32_t* src,
                          const double* buf, double* out, const int32_t* dst) {
    const int spl = kScatterSPL[variant];
    const int64_t per = ((nseg + kTile * spl - 1) / (kTile * spl) + kXCD - 1) / kXCD;
    const dim3 gr((unsigned)(per * kXCD));
    const double* nul = nullptr;
    switch (variant) {
    case 0: hipLaunchKernelGGL((scatter_kernel<false, CODED, 1, 1024, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 1: hipLaunchKernelGGL((scatter_kernel<false, CODED, 2, 1024, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 2: hipLaunchKernelGGL((scatter_kernel<false, CODED, 2, 2048, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 3: hipLaunchKernelGGL((scatter_kernel<false, CODED, 3, 2048, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 4: hipLaunchKernelGGL((scatter_kernel<false, CODED, 4, 1024, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 5: hipLaunchKernelGGL((scatter_kernel<false, CODED, 4, 2048, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 6: hipLaunchKernelGGL((scatter_kernel<false, CODED, 6, 2048, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 7: hipLaunchKernelGGL((scatter_kernel<false, CODED, 8, 2048, false, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 8: hipLaunchKernelGGL((scatter_kernel<false, CODED, 1, 1024, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 9: hipLaunchKernelGGL((scatter_kernel<false, CODED, 2, 1024, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 10: hipLaunchKernelGGL((scatter_kernel<false, CODED, 2, 2048, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 11: hipLaunchKernelGGL((scatter_kernel<false, CODED, 3, 2048, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 12: hipLaunchKernelGGL((scatter_kernel<false, CODED, 4, 1024, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 13: hipLaunchKernelGGL((scatter_kernel<false, CODED, 4, 2048, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    case 14: hipLaunchKernelGGL((scatter_kernel<false, CODED, 6, 2048, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    default: hipLaunchKernelGGL((scatter_kernel<false, CODED, 8, 2048, true, IDX>), gr, dim3(kTile), 0, s, nseg, per, seg, src, buf, nul, out, dst); break;
    }
}
// Launch one of the tuned shapes of the nz scatter; dst (optional): write segment k to out[dst[k]].
template <bool CODED>
void launch_scatter(int variant, hipStream_t s, int64_t nseg, const int32_t* seg, const uint32_t* src, const double* buf,
                    double* out, const int32_t* dst = nullptr) {
    if (dst) launch_scatter_shape<CODED, true>(variant, s, nseg, seg, src, buf, out, dst);
    else launch_scatter_shape<CODED, false>(variant, s, nseg, seg, src, buf, out, nullptr);
}

// Long segments (many elements share a dof, e.g. a U-node loading a whole stretch of
// beams): one workgroup per segment. Lane t sums a contiguous sub-range sequentially, then
// lane 0 adds the kTile partials in lane order — a fixed, deterministic order.
__global__ __launch_bounds__(kTile) void scatter_long_kernel(int64_t nseg, const int32_t* __restrict__ seg,
                                                             const uint32_t* __restrict__ src,
                                                             const double* __restrict__ buf, double* __restrict__ out) {
    __shared__ double part[kTile];
    const int64_t k = blockIdx.x;
    if (k >= nseg) return;
    const int t = threadIdx.x;
    const int32_t p0 = seg[k], p1 = seg[k + 1];
    const int32_t len = p1 - p0;
    const int32_t per = (len + kTile - 1) / kTile;
    const int32_t a = p0 + t * per, b = (a + per < p1 ? a + per : p1);
    double acc = 0.;
    for (int32_t q = a; q < b; ++q) acc += buf[src[q]];
    part[t] = acc;
    __syncthreads();
    if (t == 0) {
        double s = 0.;
        for (int i = 0; i < kTile; ++i) s += part[i];
        out[k] = s;
    }
}

__global__ void fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// y = alpha · s ⊙ x (x and y may alias)
__global__ __launch_bounds__(kBlock) void scaled_copy_kernel(int64_t n, double alpha, const double* __restrict__ sc,
                                                             const double* x, double* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = alpha * (sc[i] * x[i]);
}

// y[idx[i]] = val[i]
__global__ __launch_bounds__(kBlock) void sparse_set_kernel(int64_t n, const int64_t* __restrict__ idx,
                                                            const double* __restrict__ val, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[idx[i]] = val[i];
}
// Plain streaming copy of n doubles: the reference kernel for HBM counter calibration.
__global__ __launch_bounds__(kBlock) void copy_kernel(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = a[i];
}

// ---- pattern construction ----------------------------------------------------
// Pairs of one group in traversal order (iele, jeledof, ieledof), key = J·nX + I,
// value = coded source of the contribution.
__global__ __launch_bounds__(kBlock) void gen_pairs_kernel(int64_t nele, int nX, const int32_t* idx, int64_t nXdof,
                                                           int64_t kbase, int compact, unsigned long long* keys,
                                                           uint32_t* vals) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)nX * nX;
    if (q >= nele * per) return;
    const int64_t e = q / per;
    const int r = (int)(q - e * per);
    const int j = r / nX, i = r - j * nX;
    const unsigned long long J = (unsigned long long)idx[(int64_t)j * nele + e];
    const unsigned long long I = (unsigned long long)idx[(int64_t)i * nele + e];
    keys[q] = J * (unsigned long long)nXdof + I;
    if (compact) {
        double m;
        const int slot = mh::compact_slot(i, j, &m);
        const uint32_t code = m == 1. ? 0u : m == -1. ? 1u : m == 0.5 ? 2u : m == -0.5 ? 3u : m == 0.25 ? 4u : 5u;
        vals[q] = (code << 29) | (uint32_t)(kbase + e * mh::kCompact + slot);
    } else {
        vals[q] = (uint32_t)(kbase + e * per + i + (int64_t)nX * j);
    }
}
__global__ __launch_bounds__(kBlock) void gen_rpairs_kernel(int64_t nele, int nX, const int32_t* idx, int64_t rbase,
                                                            unsigned int* keys, uint32_t* vals) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nele * nX) return;
    const int64_t e = q / nX;
    const int i = (int)(q - e * nX);
    keys[q] = (unsigned int)idx[(int64_t)i * nele + e];
    vals[q] = (uint32_t)(rbase + (int64_t)i * nele + e);
}
template <class K>
__global__ __launch_bounds__(kBlock) void head_kernel(int64_t n, const K* keys, int32_t* head) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    head[p] = (p == 0 || keys[p] != keys[p - 1]) ? 1 : 0;
}
// For each segment head: seg[id] = p, and (row, col) of the nz from the key.
__global__ __launch_bounds__(kBlock) void seg_kernel(int64_t n, const unsigned long long* keys, const int32_t* head,
                                                     const int32_t* id, int64_t nXdof, int32_t* seg, int32_t* rowval,
                                                     int32_t* colk) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !head[p]) return;
    const int32_t k = id[p] - 1;
    seg[k] = (int32_t)p;
    rowval[k] = (int32_t)(keys[p] % (unsigned long long)nXdof);
    colk[k] = (int32_t)(keys[p] / (unsigned long long)nXdof);
}
// ptr[c] = first segment whose column is >= c  (c = 0..ncol), given sorted col[0..nseg)
__global__ __launch_bounds__(kBlock) void gapfill_kernel(int64_t nseg, const int32_t* col, int64_t ncol, int32_t* ptr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nseg) return;
    const int64_t cprev = k == 0 ? -1 : col[k - 1];
    const int64_t c = k == nseg ? ncol : col[k];
    for (int64_t cc = cprev + 1; cc <= c; ++cc) ptr[cc] = (int32_t)k;
}
__global__ __launch_bounds__(kBlock) void rseg_kernel(int64_t n, const unsigned int* keys, int64_t ndof, int32_t* seg) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    const int64_t kprev = p == 0 ? -1 : (int64_t)keys[p - 1];
    const int64_t k = p == n ? ndof : (int64_t)keys[p];
    for (int64_t cc = kprev + 1; cc <= k; ++cc) seg[cc] = (int32_t)p;
}
// Pairs of a rectangular class block (rows: nI element dofs through idxI, columns: nJ
// through idxJ) in traversal order (iele, jeledof, ieledof); key = J·nrowdof + I, value =
// the element-block entry (i + nI·j), or of the transposed block (j + nJ·i).
__global__ __launch_bounds__(kBlock) void gen_pairs_rect_kernel(int64_t nele, int nI, const int32_t* idxI, int nJ,
                                                                const int32_t* idxJ, int64_t nrowdof, int64_t kbase,
                                                                int transpose, unsigned long long* keys, uint32_t* vals) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)nI * nJ;
    if (q >= nele * per) return;
    const int64_t e = q / per;
    const int r = (int)(q - e * per);
    const int j = r / nI, i = r - j * nI;
    const unsigned long long J = (unsigned long long)idxJ[(int64_t)j * nele + e];
    const unsigned long long I = (unsigned long long)idxI[(int64_t)i * nele + e];
    keys[q] = J * (unsigned long long)nrowdof + I;
    if (vals) vals[q] = (uint32_t)(kbase + e * per + (transpose ? (j + (int64_t)nJ * i) : (i + (int64_t)nI * j)));
}
// asm[2] dump: nz number (1-based) of every element entry, found in the CSC pattern.
__global__ __launch_bounds__(kBlock) void asm_dump_kernel(int64_t nele, int nX, const int32_t* idx, const int32_t* colptr,
                                                          const int32_t* rowval, int32_t* out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)nX * nX;
    if (q >= nele * per) return;
    const int64_t e = q / per;
    const int r = (int)(q - e * per);
    const int j = r / nX, i = r - j * nX;
    const int32_t J = idx[(int64_t)j * nele + e], I = idx[(int64_t)i * nele + e];
    int32_t lo = colptr[J], hi = colptr[J + 1];
    while (lo < hi) { int32_t mid = lo + (hi - lo) / 2; if (rowval[mid] < I) lo = mid + 1; else hi = mid; }
    out[q] = lo + 1;
}

inline unsigned grid(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int check_launch(Ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { c->err = std::string(what) + ": " + hipGetErrorString(e); return MH_ERR_HIP; }
    return MH_OK;
}

}  // namespace

// Products of dof scales per compact slot (scale.X of node 1 and node 2 are equal per field).
void beam_scales(const double* sX, BeamArgs& a) {
    for (int i = 0; i < 12; ++i) a.sR[i] = sX[i];
    for (int aa = 0; aa < 3; ++aa)
        for (int bb = aa; bb < 3; ++bb) a.sA[mh::sym3(aa, bb)] = sX[aa] * sX[bb];
    for (int aa = 0; aa < 3; ++aa)
        for (int r = 0; r < 6; ++r) a.sT[6 * aa + r] = sX[aa] * sX[r < 3 ? 3 + r : 6 + r];
    for (int r = 0; r < 6; ++r)
        for (int q = r; q < 6; ++q) a.sS[mh::sym6(r, q)] = sX[r < 3 ? 3 + r : 6 + r] * sX[q < 3 ? 3 + q : 6 + q];
}

// The static beam kernel's arguments for group g (compact tangent and R into the group's slots).
BeamArgs static_beam_args(Ctx* c, const Group& g, unsigned long long gtag) {
    BeamArgs a;
    a.nele = g.nele; a.idx = g.d_idx; a.uidx = g.d_uidx; a.nodes = g.d_nodes; a.xyz = c->d_xyz; a.o2 = g.d_o2;
    a.sec = g.d_sec; a.X = c->d_X[0]; a.U = c->d_U;
    beam_scales(g.scaleX, a);
    a.K = c->d_Kbuf + g.kbase; a.R = c->d_Rbuf + g.rbase; a.flag = c->d_flag; a.gtag = gtag;
    return a;
}

// Value pass then direction pass of one dynamic beam group.
template <int ND, bool UDOF>
void launch_dyn(BeamDynArgs& a, bool step, hipStream_t s) {
    {
        const int64_t ntile = (a.nele + kBeamBlock - 1) / kBeamBlock;
        a.per_xcd = (ntile + kXCD - 1) / kXCD;
        hipLaunchKernelGGL((beam_dyn_value_kernel<ND, UDOF>), dim3((unsigned)(a.per_xcd * kXCD)), dim3(kBeamBlock), 0, s, a);
    }
    const int NE = step ? kDirElemsW<true> : kDirElemsW<false>;
    const int64_t ntile = (a.nele + NE - 1) / NE;
    a.per_xcd = (ntile + kXCD - 1) / kXCD;
    const dim3 gr((unsigned)(a.per_xcd * kXCD));
    if (step) hipLaunchKernelGGL((beam_dyn_dir_kernel<ND, UDOF, true>), gr, dim3(kDirWBlock), 0, s, a);
    else hipLaunchKernelGGL((beam_dyn_dir_kernel<ND, UDOF, false>), gr, dim3(kDirWBlock), 0, s, a);
}
int launch_dyn_beams(Ctx* c, const Group& g, int OX, bool step, const double* c6, unsigned long long gtag) {
    BeamDynArgs a;
    a.nele = g.nele; a.idx = g.d_idx; a.uidx = g.d_uidx; a.nodes = g.d_nodes; a.xyz = c->d_xyz; a.o2 = g.d_o2;
    a.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) a.X[d] = c->d_X[d];
    a.U = c->d_U;
    BeamArgs b;
    beam_scales(g.scaleX, b);
    for (int i = 0; i < 6; ++i) a.sA[i] = b.sA[i];
    for (int i = 0; i < 18; ++i) a.sT[i] = b.sT[i];
    for (int i = 0; i < 21; ++i) a.sS[i] = b.sS[i];
    for (int i = 0; i < 12; ++i) { a.sR[i] = b.sR[i]; a.scale[i] = g.scaleX[i]; }
    for (int i = 0; i < 6; ++i) a.c[i] = c6[i];
    a.dc = c->d_c6;
    a.Kc = c->d_Kc + g.kcbase; a.K = c->d_Kbuf + g.kbase; a.R = c->d_Rbuf + g.rbase;
    a.Kb = c->d_Kb + g.kcbase / mh::kCompact * mh::kDynBase;
    a.dR = step ? c->d_dRbuf + g.rbase : nullptr;
    a.flag = c->d_flag; a.gtag = gtag;
    const bool u = g.kind == MH_KIND_BEAM_U;
    if (OX == 1) { if (u) launch_dyn<2, true>(a, step, c->stream); else launch_dyn<2, false>(a, step, c->stream); }
    else { if (u) launch_dyn<3, true>(a, step, c->stream); else launch_dyn<3, false>(a, step, c->stream); }
    return check_launch(c, "dynamic beam kernels");
}

template <int ND, bool UDOF>
void launch_bar(BarArgs& a, bool step, hipStream_t s) {
    const int64_t ntile = (a.nele + kBarBlock - 1) / kBarBlock;
    a.per_xcd = (ntile + kXCD - 1) / kXCD;
    const dim3 gr((unsigned)(a.per_xcd * kXCD));
    if (step) hipLaunchKernelGGL((bar_kernel<ND, UDOF, true>), gr, dim3(kBarBlock), 0, s, a);
    else hipLaunchKernelGGL((bar_kernel<ND, UDOF, false>), gr, dim3(kBarBlock), 0, s, a);
}
int launch_bars(Ctx* c, const Group& g, int OX, bool step, const double* c6, unsigned long long gtag) {
    BarArgs a;
    a.nele = g.nele; a.idx = g.d_idx; a.uidx = g.d_uidx; a.nodes = g.d_nodes; a.xyz = c->d_xyz; a.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) a.X[d] = c->d_X[d];
    a.U = c->d_U;
    a.t = c->t;
    for (int i = 0; i < 6; ++i) { a.scale[i] = g.scaleX[i]; a.c[i] = (OX > 0 && c6) ? c6[i] : 0.; }
    a.dc = c->d_c6;
    a.K = c->d_Kbuf + g.kbase; a.R = c->d_Rbuf + g.rbase; a.dR = step ? c->d_dRbuf + g.rbase : nullptr;
    a.flag = c->d_flag; a.gtag = gtag;
    const bool u = g.kind == MH_KIND_BAR_U;
    if (OX == 0) { if (u) launch_bar<1, true>(a, false, c->stream); else launch_bar<1, false>(a, false, c->stream); }
    else if (OX == 1) { if (u) launch_bar<2, true>(a, step, c->stream); else launch_bar<2, false>(a, step, c->stream); }
    else { if (u) launch_bar<3, true>(a, step, c->stream); else launch_bar<3, false>(a, step, c->stream); }
    return check_launch(c, "bar kernels");
}

// Adjoint of one group: Λ·∂R/∂X⁽ᵈ⁾ (d < nd) and Λ·∂R/∂U into the per-element buffers.
int launch_vjp(Ctx* c, const Group& g, int nd, unsigned long long gtag) {
    hipStream_t s = c->stream;
    VjpArgs a;
    a.nele = g.nele; a.nX = g.nX; a.kind = g.kind; a.idx = g.d_idx; a.uidx = g.d_uidx; a.nodes = g.d_nodes;
    a.xyz = c->d_xyz; a.o2 = g.d_o2; a.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) a.X[d] = c->d_X[d];
    a.U = c->d_U; a.lam = c->d_lam; a.t = c->t; a.Kc = nullptr;
    a.g = c->d_gbuf + g.rbase; a.gstride = c->nR;
    a.gu = g.nU ? c->d_gubuf + g.ubase : nullptr;
    a.flag = c->d_flag; a.gtag = gtag;
    const bool beam = g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U;
    const bool bar = g.kind == MH_KIND_BAR || g.kind == MH_KIND_BAR_U;
    const bool u = g.kind == MH_KIND_BEAM_U || g.kind == MH_KIND_BAR_U;
    if (beam && nd == 1) {
        const dim3 gr((unsigned)((g.nele + kBeamBlock - 1) / kBeamBlock));
        if (u) hipLaunchKernelGGL((beam_vjp_static_kernel<true>), gr, dim3(kBeamBlock), 0, s, a);
        else hipLaunchKernelGGL((beam_vjp_static_kernel<false>), gr, dim3(kBeamBlock), 0, s, a);
    } else if (beam) {
        // value pass with unit scales: the unscaled compact static part of the frozen-load Π
        BeamDynArgs b;
        b.nele = g.nele; b.idx = g.d_idx; b.uidx = g.d_uidx; b.nodes = g.d_nodes; b.xyz = c->d_xyz; b.o2 = g.d_o2;
        b.sec = g.d_sec;
        for (int d = 0; d < 3; ++d) b.X[d] = c->d_X[d];
        b.U = c->d_U;
        for (int i = 0; i < 6; ++i) b.sA[i] = 1.;
        for (int i = 0; i < 18; ++i) b.sT[i] = 1.;
        for (int i = 0; i < 21; ++i) b.sS[i] = 1.;
        for (int i = 0; i < 12; ++i) { b.sR[i] = 1.; b.scale[i] = 1.; }
        for (int i = 0; i < 6; ++i) b.c[i] = 0.;
        b.Kc = c->d_Kc + g.kcbase; b.K = nullptr; b.R = nullptr; b.dR = nullptr;
        b.Kb = c->d_Kb + g.kcbase / mh::kCompact * mh::kDynBase;
        b.flag = c->d_flag; b.gtag = gtag;
        const int64_t ntile = (g.nele + kBeamBlock - 1) / kBeamBlock;
        b.per_xcd = (ntile + kXCD - 1) / kXCD;
        const dim3 grv((unsigned)(b.per_xcd * kXCD));
        if (nd == 2) {
            if (u) hipLaunchKernelGGL((beam_dyn_value_kernel<2, true>), grv, dim3(kBeamBlock), 0, s, b);
            else hipLaunchKernelGGL((beam_dyn_value_kernel<2, false>), grv, dim3(kBeamBlock), 0, s, b);
        } else {
            if (u) hipLaunchKernelGGL((beam_dyn_value_kernel<3, true>), grv, dim3(kBeamBlock), 0, s, b);
            else hipLaunchKernelGGL((beam_dyn_value_kernel<3, false>), grv, dim3(kBeamBlock), 0, s, b);
        }
        a.Kc = b.Kc;
        a.Kb = b.Kb;
        a.yb = c->d_ybuf + g.kcbase / mh::kCompact * 15;
        {
            const dim3 gy((unsigned)((g.nele + kBlock - 1) / kBlock));
            if (nd == 2) {
                if (u) hipLaunchKernelGGL((beam_dyn_y_kernel<2, true>), gy, dim3(kBlock), 0, s, a);
                else hipLaunchKernelGGL((beam_dyn_y_kernel<2, false>), gy, dim3(kBlock), 0, s, a);
            } else {
                if (u) hipLaunchKernelGGL((beam_dyn_y_kernel<3, true>), gy, dim3(kBlock), 0, s, a);
                else hipLaunchKernelGGL((beam_dyn_y_kernel<3, false>), gy, dim3(kBlock), 0, s, a);
            }
        }
        const int64_t nt = (g.nele + kDirElemsW<false> - 1) / kDirElemsW<false>;
        a.per_xcd = (nt + kXCD - 1) / kXCD;
        const dim3 gr((unsigned)(a.per_xcd * kXCD));
        if (nd == 2) {
            if (u) {
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<2, true, 0>), gr, dim3(kDirWBlock), 0, s, a);
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<2, true, 1>), gr, dim3(kDirWBlock), 0, s, a);
            } else {
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<2, false, 0>), gr, dim3(kDirWBlock), 0, s, a);
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<2, false, 1>), gr, dim3(kDirWBlock), 0, s, a);
            }
        } else {
            if (u) {
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<3, true, 0>), gr, dim3(kDirWBlock), 0, s, a);
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<3, true, 1>), gr, dim3(kDirWBlock), 0, s, a);
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<3, true, 2>), gr, dim3(kDirWBlock), 0, s, a);
            } else {
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<3, false, 0>), gr, dim3(kDirWBlock), 0, s, a);
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<3, false, 1>), gr, dim3(kDirWBlock), 0, s, a);
                hipLaunchKernelGGL((beam_dyn_vjp_kernel<3, false, 2>), gr, dim3(kDirWBlock), 0, s, a);
            }
        }
    } else if (bar) {
        const dim3 gr((unsigned)((g.nele + kBarBlock - 1) / kBarBlock));
        if (nd == 1) { if (u) hipLaunchKernelGGL((bar_vjp_kernel<1, true>), gr, dim3(kBarBlock), 0, s, a);
                       else hipLaunchKernelGGL((bar_vjp_kernel<1, false>), gr, dim3(kBarBlock), 0, s, a); }
        else if (nd == 2) { if (u) hipLaunchKernelGGL((bar_vjp_kernel<2, true>), gr, dim3(kBarBlock), 0, s, a);
                            else hipLaunchKernelGGL((bar_vjp_kernel<2, false>), gr, dim3(kBarBlock), 0, s, a); }
        else { if (u) hipLaunchKernelGGL((bar_vjp_kernel<3, true>), gr, dim3(kBarBlock), 0, s, a);
               else hipLaunchKernelGGL((bar_vjp_kernel<3, false>), gr, dim3(kBarBlock), 0, s, a); }
    } else {
        hipLaunchKernelGGL(point_vjp_kernel, dim3(grid(g.nele)), dim3(kBlock), 0, s, a, nd);
    }
    if (beam && g.amask)
        hipLaunchKernelGGL(beam_vjp_a_kernel<false>, dim3((unsigned)((g.nele + kBeamBlock - 1) / kBeamBlock)), dim3(kBeamBlock),
                           0, s, a, g.amask, c->d_gabuf + g.abase);
    return check_launch(c, "adjoint kernels");
}

// The scaled sections of the EulerBeam3DA groups at the device A (after mh_set_A and
// mh_group_set_adofs).
int apply_ascale(Ctx* c) {
    for (auto& g : c->groups) {
        if (!g.amask || g.nele == 0) continue;
        hipLaunchKernelGGL(scale_section_kernel, dim3(grid(g.nele)), dim3(kBlock), 0, c->stream, g.nele, g.d_aidx, c->d_A,
                           g.amask, g.d_sec_base, g.d_sec);
    }
    return check_launch(c, "section scaling");
}

// ============================================================================ C ABI
struct mh_ctx : Ctx {};

// A copy between host and device memory (or device to device) in the context's stream
// order, returning when it is complete. The context's stream is non-blocking, so copies on
// the null stream (hipMemcpy) would not be ordered with its kernels.
static hipError_t copy_sync(Ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}
// Zero device memory at set-up, in the context's stream order (never inside a captured graph).
static hipError_t zero_sync(Ctx* c, void* dst, size_t bytes) {
    const hipError_t e = hipMemsetAsync(dst, 0, bytes, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

static int upload_idx(Ctx* c, int32_t** dst, const int64_t* src, int n, int64_t nele, int64_t ndof, bool allow_zero) {
    std::vector<int32_t> h((size_t)n * nele);
    for (int64_t e = 0; e < nele; ++e)
        for (int k = 0; k < n; ++k) {
            int64_t v = src[e * n + k];
            if (v < (allow_zero ? 0 : 1) || v > ndof) { c->err = "dof index out of range"; return MH_ERR_ARG; }
            h[(size_t)k * nele + e] = (int32_t)(v - 1);
        }
    HIPCHK(c, hipMalloc(dst, sizeof(int32_t) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HIPCHK(c, copy_sync(c, *dst, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
    return MH_OK;
}
template <class T>
static int upload_soa(Ctx* c, T** dst, const T* src, int n, int64_t nele) {
    std::vector<T> h((size_t)n * nele);
    for (int64_t e = 0; e < nele; ++e)
        for (int k = 0; k < n; ++k) h[(size_t)k * nele + e] = src[e * n + k];
    HIPCHK(c, hipMalloc(dst, sizeof(T) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HIPCHK(c, copy_sync(c, *dst, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return MH_OK;
}


// The rhs map (asm[1] for the X dofs): contributions of every element dof into Lλ, in
// traversal order, stably sorted by dof (Assemble.jl:340-357 with add_value!'s order).
static int build_rmap(Ctx* c) {
    hipStream_t s = c->stream;
    const int64_t nR = c->nR, nX = c->nX;
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    HIPCHK(c, hipMalloc(&c->d_rseg, sizeof(int32_t) * (nX + 1)));
    unsigned int *rkin = nullptr, *rkout = nullptr;
    uint32_t* rvin = nullptr;
    HIPCHK(c, hipMalloc(&rkin, sizeof(unsigned int) * std::max<int64_t>(nR, 1)));
    HIPCHK(c, hipMalloc(&rkout, sizeof(unsigned int) * std::max<int64_t>(nR, 1)));
    HIPCHK(c, hipMalloc(&rvin, sizeof(uint32_t) * std::max<int64_t>(nR, 1)));
    HIPCHK(c, hipMalloc(&c->d_rsrc, sizeof(uint32_t) * std::max<int64_t>(nR, 1)));
    for (auto& g : c->groups) {
        if (g.nX == 0 || g.nele == 0) continue;
        int64_t np = g.nele * g.nX;
        hipLaunchKernelGGL(gen_rpairs_kernel, dim3(grid(np)), dim3(kBlock), 0, s, g.nele, g.nX, g.d_idx, g.rbase,
                           rkin + g.rbase, rvin + g.rbase);
    }
    int rbit = 1;
    while (rbit < 32 && (unsigned long long)nX > (1ull << rbit)) ++rbit;
    tmp_bytes = 0;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, rkin, rkout, rvin, c->d_rsrc, (int)nR, 0, rbit, s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, rkin, rkout, rvin, c->d_rsrc, (int)nR, 0, rbit, s));
    dfree(tmp);
    hipLaunchKernelGGL(rseg_kernel, dim3(grid(nR + 1)), dim3(kBlock), 0, s, nR, rkout, nX, c->d_rseg);
    if (int r = check_launch(c, "rhs pattern")) return r;
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(rkin); dfree(rkout); dfree(rvin);
    return MH_OK;
}

// One class-pair block of AssemblyDirect: asmmat!(asm[α,β], asm[α], asm[β]) on the device
// (Assemble.jl:373-448) — pairs in traversal order, stable radix sort on (column, row),
// run-length numbering; src maps every pair to its element-block entry (srcT: to the
// transposed entry, same order).
struct PairSpec { int64_t nele; int nI; const int32_t* idxI; int nJ; const int32_t* idxJ; int64_t kbase; };
static int build_dmap(Ctx* c, const std::vector<PairSpec>& ps, int64_t nrow, int64_t ncol, bool want_src, bool want_srcT,
                      DMap& M) {
    hipStream_t s = c->stream;
    int64_t np = 0;
    for (auto& p : ps) np += p.nele * p.nI * p.nJ;
    M.nrow = nrow; M.ncol = ncol; M.npair = np; M.nnz = 0;
    HIPCHK(c, hipMalloc(&M.colptr, sizeof(int32_t) * (ncol + 1)));
    if (np == 0) {
        HIPCHK(c, hipMemsetAsync(M.colptr, 0, sizeof(int32_t) * (ncol + 1), s));
        HIPCHK(c, hipStreamSynchronize(s));
        return MH_OK;
    }
    if (np >= INT32_MAX || np >= (int64_t)UINT32_MAX) { c->err = "too many contributions for 32-bit maps"; return MH_ERR_SIZE; }
    unsigned long long *kin = nullptr, *kout = nullptr;
    uint32_t* vin = nullptr;
    int32_t *head = nullptr, *id = nullptr, *colk = nullptr;
    HIPCHK(c, hipMalloc(&kin, sizeof(unsigned long long) * np));
    HIPCHK(c, hipMalloc(&kout, sizeof(unsigned long long) * np));
    HIPCHK(c, hipMalloc(&vin, sizeof(uint32_t) * np));
    int end_bit = 1;
    while (end_bit < 64 && ((unsigned long long)nrow * (unsigned long long)ncol) > (1ull << end_bit)) ++end_bit;
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    // one sort per wanted source map (stable, identical keys: identical order); at least one
    // sort always runs, for the sorted keys
    for (int pass = 0; pass < 2; ++pass) {
        const bool tr = pass == 1;
        if (tr ? !want_srcT : (!want_src && want_srcT)) continue;
        uint32_t* dst = nullptr;
        HIPCHK(c, hipMalloc(&dst, sizeof(uint32_t) * np));
        int64_t off = 0;
        for (auto& p : ps) {
            const int64_t n = p.nele * p.nI * p.nJ;
            if (n == 0) continue;
            hipLaunchKernelGGL(gen_pairs_rect_kernel, dim3(grid(n)), dim3(kBlock), 0, s, p.nele, p.nI, p.idxI, p.nJ, p.idxJ,
                               nrow, p.kbase, tr ? 1 : 0, kin + off, vin + off);
            off += n;
        }
        if (int r = check_launch(c, "direct pairs")) return r;
        tmp_bytes = 0;
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, vin, dst, (int)np, 0, end_bit, s));
        HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, dst, (int)np, 0, end_bit, s));
        HIPCHK(c, hipStreamSynchronize(s));
        dfree(tmp);
        if (tr) M.srcT = dst;
        else if (want_src) M.src = dst;
        else dfree(dst);
    }
    HIPCHK(c, hipMalloc(&head, sizeof(int32_t) * np));
    HIPCHK(c, hipMalloc(&id, sizeof(int32_t) * np));
    hipLaunchKernelGGL(head_kernel<unsigned long long>, dim3(grid(np)), dim3(kBlock), 0, s, np, kout, head);
    tmp_bytes = 0;
    HIPCHK(c, hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, head, id, (int)np, s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, head, id, (int)np, s));
    dfree(tmp);
    int32_t nnz = 0;
    HIPCHK(c, hipMemcpyAsync(&nnz, id + np - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    M.nnz = nnz;
    HIPCHK(c, hipMalloc(&M.rowval, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    HIPCHK(c, hipMalloc(&M.seg, sizeof(int32_t) * (nnz + 1)));
    HIPCHK(c, hipMalloc(&colk, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    hipLaunchKernelGGL(seg_kernel, dim3(grid(np)), dim3(kBlock), 0, s, np, kout, head, id, nrow, M.seg, M.rowval, colk);
    int32_t np32 = (int32_t)np;
    HIPCHK(c, hipMemcpyAsync(M.seg + nnz, &np32, sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gapfill_kernel, dim3(grid(nnz + 1)), dim3(kBlock), 0, s, (int64_t)nnz, colk, ncol, M.colptr);
    if (int r = check_launch(c, "direct pattern")) return r;
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(kin); dfree(kout); dfree(vin); dfree(head); dfree(id); dfree(colk);
    return MH_OK;
}

// Value pass then the direct-blocks pass of one beam group.
template <int ND>
static void launch_direct_beam(bool u, dim3 grv, dim3 gr, hipStream_t s, const BeamDynArgs& b, const DirectArgs& a) {
    if (u) {
        hipLaunchKernelGGL((beam_dyn_value_kernel<ND, true>), grv, dim3(kBeamBlock), 0, s, b);
        hipLaunchKernelGGL((beam_direct_kernel<ND, true>), gr, dim3(kDirBlock), 0, s, a);
    } else {
        hipLaunchKernelGGL((beam_dyn_value_kernel<ND, false>), grv, dim3(kBeamBlock), 0, s, b);
        hipLaunchKernelGGL((beam_direct_kernel<ND, false>), gr, dim3(kDirBlock), 0, s, a);
    }
}

// Element blocks of one group for a DirectXUA step.
static int launch_direct(Ctx* c, const Group& g, int nd, double dt, unsigned long long gtag) {
    hipStream_t s = c->stream;
    DirectArgs a;
    a.nele = g.nele; a.nX = g.nX; a.kind = g.kind; a.idx = g.d_idx; a.uidx = g.d_uidx; a.nodes = g.d_nodes;
    a.xyz = c->d_xyz; a.o2 = g.d_o2; a.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) a.X[d] = c->d_X[d];
    a.U = c->d_U; a.lam = c->d_lam; a.vals = g.d_vals; a.t = c->t; a.dt = dt;
    for (int i = 0; i < 12; ++i) { a.sX[i] = g.scaleX[i]; a.sL[i] = g.scaleL[i]; }
    for (int i = 0; i < 3; ++i) a.sU[i] = g.scaleU[i];
    a.Kc = nullptr;
    a.Kd = c->d_Kd + g.kdbase; a.nKd = c->nKd;
    a.KU = g.nU ? c->d_KU + g.kubase : nullptr;
    a.R = c->d_Rd + g.rbase; a.L1x = c->d_L1x + g.rbase;
    a.flag = c->d_flag; a.gtag = gtag;
    const bool beam = g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U;
    const bool bar = g.kind == MH_KIND_BAR || g.kind == MH_KIND_BAR_U;
    const bool u = g.kind == MH_KIND_BEAM_U || g.kind == MH_KIND_BAR_U;
    if (beam) {
        // value pass with unit scales (unscaled compact static part), R × Δt
        BeamDynArgs b;
        b.nele = g.nele; b.idx = g.d_idx; b.uidx = g.d_uidx; b.nodes = g.d_nodes; b.xyz = c->d_xyz; b.o2 = g.d_o2;
        b.sec = g.d_sec;
        for (int d = 0; d < 3; ++d) b.X[d] = c->d_X[d];
        b.U = c->d_U;
        for (int i = 0; i < 6; ++i) b.sA[i] = 1.;
        for (int i = 0; i < 18; ++i) b.sT[i] = 1.;
        for (int i = 0; i < 21; ++i) b.sS[i] = 1.;
        for (int i = 0; i < 12; ++i) { b.sR[i] = dt; b.scale[i] = 1.; }
        for (int i = 0; i < 6; ++i) b.c[i] = 0.;
        b.Kc = c->d_Kc + g.kcbase; b.K = nullptr; b.R = c->d_Rd + g.rbase; b.dR = nullptr;
        b.flag = c->d_flag; b.gtag = gtag;
        const int64_t ntile = (g.nele + kBeamBlock - 1) / kBeamBlock;
        b.per_xcd = (ntile + kXCD - 1) / kXCD;
        const dim3 grv((unsigned)(b.per_xcd * kXCD));
        a.Kc = b.Kc;
        const int64_t nt = (g.nele + kDirElems - 1) / kDirElems;
        a.per_xcd = (nt + kXCD - 1) / kXCD;
        const dim3 gr((unsigned)(a.per_xcd * kXCD));
        if (nd == 1) launch_direct_beam<1>(u, grv, gr, s, b, a);
        else if (nd == 2) launch_direct_beam<2>(u, grv, gr, s, b, a);
        else launch_direct_beam<3>(u, grv, gr, s, b, a);
    } else if (bar) {
        const dim3 gr((unsigned)((g.nele + kBarBlock - 1) / kBarBlock));
        if (nd == 1) { if (u) hipLaunchKernelGGL((bar_direct_kernel<1, true>), gr, dim3(kBarBlock), 0, s, a);
                       else hipLaunchKernelGGL((bar_direct_kernel<1, false>), gr, dim3(kBarBlock), 0, s, a); }
        else if (nd == 2) { if (u) hipLaunchKernelGGL((bar_direct_kernel<2, true>), gr, dim3(kBarBlock), 0, s, a);
                            else hipLaunchKernelGGL((bar_direct_kernel<2, false>), gr, dim3(kBarBlock), 0, s, a); }
        else { if (u) hipLaunchKernelGGL((bar_direct_kernel<3, true>), gr, dim3(kBarBlock), 0, s, a);
               else hipLaunchKernelGGL((bar_direct_kernel<3, false>), gr, dim3(kBarBlock), 0, s, a); }
    } else {
        hipLaunchKernelGGL(point_direct_kernel, dim3(grid(g.nele)), dim3(kBlock), 0, s, a, nd);
    }
    return check_launch(c, "direct element kernels");
}

// ------------------------------------------------------------------ fixed scratch
// Device memory of the entry points that allocate nothing (mh_beam_gauge_derivs_noalloc): a
// file-scope array that every GPU the library runs on holds until the process ends. Host
// code never passes the array's name to a kernel or a copy (on the host it is only a
// stand-in): it uses the per-device address that hipGetSymbolAddress gives, looked up once
// per device when the first context on that device is created.
namespace {
constexpr int64_t kGaugeNoAllocMax = 4096;                         // elements per call
constexpr size_t kScratchSelBytes = sizeof(int64_t) * kGaugeNoAllocMax;
constexpr size_t kScratchBytes = kScratchSelBytes + sizeof(double) * kGaugeOut * kGaugeNoAllocMax;   // 11.9 MB
static_assert(kScratchSelBytes % 256 == 0, "output staging must stay 256-byte aligned");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices] = {};

// The calling thread's current device must be `device`.
int scratch_address(int device, void** p) {
    *p = nullptr;
    if (device < 0 || device >= kScratchMaxDevices) return MH_OK;   // no fixed scratch: the no-alloc entries refuse
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return MH_ERR_HIP;
        g_scratch_ptr[device] = q;
    }
    *p = g_scratch_ptr[device];
    return MH_OK;
}
}  // namespace

extern "C" {

const char* mh_strerror(int code) {
    switch (code) {
    case MH_OK: return "ok";
    case MH_ERR_ARG: return "invalid argument or call order";
    case MH_ERR_HIP: return "HIP runtime error";
    case MH_ERR_NAN: return "element returned NaN in R or its derivatives";
    case MH_ERR_ORIENT: return "orient2 is nearly parallel to a beam element";
    case MH_ERR_SIZE: return "model exceeds the 32-bit index range of the device maps";
    case MH_ERR_SOLVE: return "linear solve failed";
    }
    return "unknown error";
}

const char* mh_last_error(const mh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mh_ctx_create(int device, mh_ctx** out) {
    if (!out) return MH_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return MH_ERR_HIP;
    mh_ctx* c = new mh_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c; return MH_ERR_HIP;
    }
    for (auto& e : c->ev) if (hipEventCreate(&e) != hipSuccess) { delete c; return MH_ERR_HIP; }
    if (hipMalloc(&c->d_flag, sizeof(unsigned long long)) != hipSuccess) { delete c; return MH_ERR_HIP; }
    if (hipHostMalloc((void**)&c->h_flag, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(c->d_flag); delete c; return MH_ERR_HIP;
    }
    if (hipMalloc(&c->d_c6, sizeof(double) * 6) != hipSuccess ||
        hipHostMalloc((void**)&c->h_c6, sizeof(double) * 6, hipHostMallocDefault) != hipSuccess ||
        scratch_address(device, &c->d_scratch) != MH_OK) {
        if (c->h_c6) (void)hipHostFree(c->h_c6);
        if (c->d_c6) (void)hipFree(c->d_c6);
        (void)hipHostFree(c->h_flag); (void)hipFree(c->d_flag); delete c; return MH_ERR_HIP;
    }
    *out = c;
    return MH_OK;
}

int mh_ctx_destroy(mh_ctx* c) {
    if (!c) return MH_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_model(c);
    dfree(c->d_flag);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    dfree(c->d_c6);
    if (c->h_c6) (void)hipHostFree(c->h_c6);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->blas) (void)rocblas_destroy_handle(c->blas);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MH_OK;
}

int mh_model_begin(mh_ctx* c, int64_t nX, int64_t nU, int64_t nA, int64_t nnod, const double* xyz) {
    if (!c || nX < 0 || nU < 0 || nA < 0 || nnod < 0 || (nnod > 0 && !xyz)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_model(c);
    if (nX >= INT32_MAX) { c->err = "too many X dofs"; return MH_ERR_SIZE; }
    c->nX = nX; c->nU = nU; c->nA = nA; c->nnod = nnod;
    if (nnod) {
        HIPCHK(c, hipMalloc(&c->d_xyz, sizeof(double) * 3 * nnod));
        HIPCHK(c, copy_sync(c, c->d_xyz, xyz, sizeof(double) * 3 * nnod, hipMemcpyHostToDevice));
    }
    for (int d = 0; d < 3; ++d) {
        HIPCHK(c, hipMalloc(&c->d_X[d], sizeof(double) * std::max<int64_t>(nX, 1)));
        HIPCHK(c, zero_sync(c, c->d_X[d], sizeof(double) * std::max<int64_t>(nX, 1)));
    }
    HIPCHK(c, hipMalloc(&c->d_U, sizeof(double) * std::max<int64_t>(nU, 1)));
    HIPCHK(c, zero_sync(c, c->d_U, sizeof(double) * std::max<int64_t>(nU, 1)));
    HIPCHK(c, hipMalloc(&c->d_A, sizeof(double) * std::max<int64_t>(nA, 1)));
    HIPCHK(c, zero_sync(c, c->d_A, sizeof(double) * std::max<int64_t>(nA, 1)));
    return MH_OK;
}

int mh_group_add(mh_ctx* c, int kind, int64_t nele, int nX, const int64_t* idxX, const double* scaleX, int nU,
                 const int64_t* idxU, const int64_t* nodes, const double* orient2, const double* mat, int* igroup) {
    if (!c || nele < 0 || nX < 0 || nX > 12 || (nele > 0 && nX > 0 && !idxX)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    free_prepared(c);
    const bool beam = kind == MH_KIND_BEAM || kind == MH_KIND_BEAM_U;
    const bool bar = kind == MH_KIND_BAR || kind == MH_KIND_BAR_U;
    if (beam && (nX != 12 || !nodes || !orient2 || !mat)) { c->err = "beam group needs 12 X dofs, nodes, orient2 and mat"; return MH_ERR_ARG; }
    if (bar && (nX != 6 || !nodes || !mat)) { c->err = "bar group needs 6 X dofs, nodes and mat"; return MH_ERR_ARG; }
    if ((kind == MH_KIND_BEAM_U || kind == MH_KIND_BAR_U) && (nU != 3 || !idxU)) {
        c->err = "Udof beam or bar group needs 3 U dofs"; return MH_ERR_ARG;
    }
    if (kind == MH_KIND_HOLD && nX != 2) { c->err = "Hold group needs 2 X dofs"; return MH_ERR_ARG; }
    if (kind == MH_KIND_DOFLOAD && nX != 1) { c->err = "DofLoad group needs 1 X dof"; return MH_ERR_ARG; }
    if (kind < MH_KIND_BEAM || kind > MH_KIND_HOST) { c->err = "unknown element kind"; return MH_ERR_ARG; }
    if (kind == MH_KIND_HOST && (nX < 1 || nX > 12)) { c->err = "host-evaluated group needs 1..12 X dofs"; return MH_ERR_ARG; }
    Group g;
    g.kind = kind; g.nele = nele; g.nX = nX; g.nU = (beam || bar) ? nU : 0;
    for (int k = 0; k < 12; ++k) g.scaleX[k] = (scaleX && k < nX) ? scaleX[k] : 1.;
    for (int k = 0; k < 12; ++k) g.scaleL[k] = g.scaleX[k];
    for (int k = 0; k < 3; ++k) g.scaleU[k] = 1.;
    if (beam)
        for (int k = 0; k < 6; ++k)
            if (g.scaleX[k] != g.scaleX[k + 6]) { c->err = "beam dof scales must be equal on both nodes"; return MH_ERR_ARG; }
    int r;
    if ((r = upload_idx(c, &g.d_idx, idxX, nX, nele, c->nX, false))) return r;
    g.h_idx.assign(idxX, idxX + (size_t)nX * nele);
    if ((kind == MH_KIND_BEAM_U || kind == MH_KIND_BAR_U) && (r = upload_idx(c, &g.d_uidx, idxU, 3, nele, c->nU, false)))
        return r;
    if (bar) {
        std::vector<int32_t> nd32((size_t)2 * nele);
        for (int64_t e = 0; e < nele; ++e)
            for (int k = 0; k < 2; ++k) {
                int64_t v = nodes[e * 2 + k];
                if (v < 1 || v > c->nnod) { c->err = "node index out of range"; return MH_ERR_ARG; }
                nd32[e * 2 + k] = (int32_t)(v - 1);
            }
        if ((r = upload_soa<int32_t>(c, &g.d_nodes, nd32.data(), 2, nele))) return r;
        if ((r = upload_soa<double>(c, &g.d_sec, mat, 10, nele))) return r;
    }
    if (beam) {
        std::vector<int64_t> nd((size_t)2 * nele);
        for (int64_t e = 0; e < nele; ++e)
            for (int k = 0; k < 2; ++k) {
                int64_t v = nodes[e * 2 + k];
                if (v < 1 || v > c->nnod) { c->err = "node index out of range"; return MH_ERR_ARG; }
                nd[e * 2 + k] = v - 1;
            }
        // orient2 check of the constructor (BeamElement.jl:132-135), on the host
        std::vector<double> xyz((size_t)3 * c->nnod);
        if (c->nnod) HIPCHK(c, copy_sync(c, xyz.data(), c->d_xyz, sizeof(double) * xyz.size(), hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < nele; ++e) {
            const double* a = &xyz[3 * nd[e * 2]];
            const double* b = &xyz[3 * nd[e * 2 + 1]];
            mh::BeamGeo G = mh::beam_geometry(mh::mk(a[0], a[1], a[2]), mh::mk(b[0], b[1], b[2]),
                                              mh::mk(orient2[3 * e], orient2[3 * e + 1], orient2[3 * e + 2]));
            if (G.bad) { c->err = "orient2 nearly parallel to beam " + std::to_string(e + 1); return MH_ERR_ORIENT; }
        }
        std::vector<int32_t> nd32(nd.size());
        for (size_t i = 0; i < nd.size(); ++i) nd32[i] = (int32_t)nd[i];
        if ((r = upload_soa<int32_t>(c, &g.d_nodes, nd32.data(), 2, nele))) return r;
        if ((r = upload_soa<double>(c, &g.d_o2, orient2, 3, nele))) return r;
        if ((r = upload_soa<double>(c, &g.d_sec, mat, 16, nele))) return r;
    }
    if (kind == MH_KIND_DOFLOAD || kind == MH_KIND_HOLD) {   // DofLoad values, prescribed-motion gaps
        HIPCHK(c, hipMalloc(&g.d_vals, sizeof(double) * std::max<int64_t>(nele, 1)));
        HIPCHK(c, zero_sync(c, g.d_vals, sizeof(double) * std::max<int64_t>(nele, 1)));
    }
    c->groups.push_back(std::move(g));
    if (igroup) *igroup = (int)c->groups.size() - 1;
    return MH_OK;
}

int mh_group_set_adofs(mh_ctx* c, int igroup, const int64_t* idxA, int mask) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size() || mask < 1 || mask > 15) return MH_ERR_ARG;
    Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_BEAM && g.kind != MH_KIND_BEAM_U) { c->err = "A-scaled stiffness needs an EulerBeam3D group"; return MH_ERR_ARG; }
    if (g.amask) { c->err = "the group's A-dofs are already set"; return MH_ERR_ARG; }
    if (g.nele > 0 && !idxA) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    free_prepared(c);
    if (int r = upload_idx(c, &g.d_aidx, idxA, 1, g.nele, c->nA, false)) return r;
    HIPCHK(c, hipMalloc(&g.d_sec_base, sizeof(double) * std::max<int64_t>(16 * g.nele, 1)));
    if (g.nele) HIPCHK(c, copy_sync(c, g.d_sec_base, g.d_sec, sizeof(double) * 16 * g.nele, hipMemcpyDeviceToDevice));
    g.amask = mask;
    if (int r = apply_ascale(c)) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}

int mh_prepare_sweepx(mh_ctx* c, int OX, int64_t* nnz_out) {
    if (!c || OX < 0 || OX > 2) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    free_prepared(c);
    // contribution counts and buffer offsets, in eletype order. Static beams store the
    // compact tangent; dynamic beams (K + a₁C + b₁M is not symmetric) the full block.
    int64_t nK = 0, nP = 0, nR = 0, nKc = 0, nRU = 0, nRA = 0;
    bool coded = false;
    for (auto& g : c->groups) {
        g.ubase = nRU;
        nRU += g.nele * g.nU;
        g.abase = nRA;
        if (g.amask) nRA += g.nele;
        const bool beam = g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U;
        g.compact = beam && OX == 0;
        coded = coded || g.compact;
        g.kper = g.compact ? mh::kCompact : (int64_t)g.nX * g.nX;
        g.kbase = nK; g.rbase = nR;
        g.kcbase = nKc;
        if (beam && OX > 0) nKc += g.nele * mh::kCompact;
        nK += g.nele * g.kper;
        nP += g.nele * g.nX * g.nX;
        nR += g.nele * g.nX;
    }
    const int64_t klim = coded ? (1ll << 29) : (int64_t)UINT32_MAX;
    if (nP >= INT32_MAX || nK >= klim || nR >= (int64_t)UINT32_MAX) {
        c->err = "more than 2^31 contributions or too many element outputs for 32-bit maps"; return MH_ERR_SIZE;
    }
    c->OX = OX; c->coded = coded; c->nRU = nRU; c->nRA = nRA;
    c->nK = nK; c->nP = nP; c->nR = nR;
    const int64_t nX = c->nX;
    hipStream_t s = c->stream;
    HIPCHK(c, hipMalloc(&c->d_Kbuf, sizeof(double) * std::max<int64_t>(nK, 1)));
    HIPCHK(c, hipMalloc(&c->d_Rbuf, sizeof(double) * std::max<int64_t>(nR, 1)));
    HIPCHK(c, hipMalloc(&c->d_dRbuf, sizeof(double) * std::max<int64_t>(nR, 1)));
    if (nKc) HIPCHK(c, hipMalloc(&c->d_Kc, sizeof(double) * nKc));
    if (nKc) HIPCHK(c, hipMalloc(&c->d_Kb, sizeof(double) * (nKc / mh::kCompact * mh::kDynBase)));
    HIPCHK(c, hipMalloc(&c->d_rhs, sizeof(double) * std::max<int64_t>(nX, 1)));
    HIPCHK(c, hipMalloc(&c->d_colptr, sizeof(int32_t) * (nX + 1)));
    // ---- matrix pairs
    unsigned long long *kin = nullptr, *kout = nullptr;
    uint32_t* vin = nullptr;
    int32_t *head = nullptr, *id = nullptr, *colk = nullptr;
    HIPCHK(c, hipMalloc(&kin, sizeof(unsigned long long) * std::max<int64_t>(nP, 1)));
    HIPCHK(c, hipMalloc(&kout, sizeof(unsigned long long) * std::max<int64_t>(nP, 1)));
    HIPCHK(c, hipMalloc(&vin, sizeof(uint32_t) * std::max<int64_t>(nP, 1)));
    HIPCHK(c, hipMalloc(&c->d_ksrc, sizeof(uint32_t) * std::max<int64_t>(nP, 1)));
    {
        int64_t pbase = 0;
        for (auto& g : c->groups) {
            int64_t np = g.nele * g.nX * g.nX;
            if (np == 0) continue;
            hipLaunchKernelGGL(gen_pairs_kernel, dim3(grid(np)), dim3(kBlock), 0, s, g.nele, g.nX, g.d_idx, nX, g.kbase,
                               (int)g.compact, kin + pbase, vin + pbase);
            pbase += np;
        }
    }
    if (int r = check_launch(c, "gen_pairs")) return r;
    int end_bit = 1;
    while (end_bit < 64 && ((unsigned long long)nX * (unsigned long long)nX) > (1ull << end_bit)) ++end_bit;
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, vin, c->d_ksrc, (int)nP, 0, end_bit, s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, c->d_ksrc, (int)nP, 0, end_bit, s));
    dfree(tmp);
    HIPCHK(c, hipMalloc(&head, sizeof(int32_t) * std::max<int64_t>(nP, 1)));
    HIPCHK(c, hipMalloc(&id, sizeof(int32_t) * std::max<int64_t>(nP, 1)));
    hipLaunchKernelGGL(head_kernel<unsigned long long>, dim3(grid(nP)), dim3(kBlock), 0, s, nP, kout, head);
    tmp_bytes = 0;
    HIPCHK(c, hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, head, id, (int)nP, s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, head, id, (int)nP, s));
    dfree(tmp);
    int32_t nnz = 0;
    if (nP > 0) HIPCHK(c, hipMemcpyAsync(&nnz, id + nP - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->nnz = nnz;
    HIPCHK(c, hipMalloc(&c->d_rowval, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    HIPCHK(c, hipMalloc(&c->d_kseg, sizeof(int32_t) * (nnz + 1)));
    HIPCHK(c, hipMalloc(&colk, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    HIPCHK(c, hipMalloc(&c->d_nzval, sizeof(double) * std::max<int64_t>(nnz, 1)));
    hipLaunchKernelGGL(seg_kernel, dim3(grid(nP)), dim3(kBlock), 0, s, nP, kout, head, id, nX, c->d_kseg, c->d_rowval, colk);
    int32_t nk32 = (int32_t)nP;
    HIPCHK(c, hipMemcpyAsync(c->d_kseg + nnz, &nk32, sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gapfill_kernel, dim3(grid(nnz + 1)), dim3(kBlock), 0, s, (int64_t)nnz, colk, nX, c->d_colptr);
    if (int r = check_launch(c, "pattern")) return r;
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(kin); dfree(kout); dfree(vin); dfree(head); dfree(id); dfree(colk);
    if (int r = build_rmap(c)) return r;
    unsigned int *rkin = nullptr, *rkout = nullptr;
    uint32_t* rvin = nullptr;
    // ---- U-dof map (gradients ∂L/∂U of the adjoint), same construction over the U class
    if (nRU > 0) {
        HIPCHK(c, hipMalloc(&rkin, sizeof(unsigned int) * nRU));
        HIPCHK(c, hipMalloc(&rkout, sizeof(unsigned int) * nRU));
        HIPCHK(c, hipMalloc(&rvin, sizeof(uint32_t) * nRU));
        HIPCHK(c, hipMalloc(&c->d_usrc, sizeof(uint32_t) * nRU));
        HIPCHK(c, hipMalloc(&c->d_useg, sizeof(int32_t) * (c->nU + 1)));
        for (auto& g : c->groups) {
            if (g.nU == 0 || g.nele == 0) continue;
            int64_t np = g.nele * g.nU;
            hipLaunchKernelGGL(gen_rpairs_kernel, dim3(grid(np)), dim3(kBlock), 0, s, g.nele, g.nU, g.d_uidx, g.ubase,
                               rkin + g.ubase, rvin + g.ubase);
        }
        int ubit = 1;
        while (ubit < 32 && (unsigned long long)c->nU > (1ull << ubit)) ++ubit;
        tmp_bytes = 0;
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, rkin, rkout, rvin, c->d_usrc, (int)nRU, 0, ubit, s));
        HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, rkin, rkout, rvin, c->d_usrc, (int)nRU, 0, ubit, s));
        dfree(tmp);
        hipLaunchKernelGGL(rseg_kernel, dim3(grid(nRU + 1)), dim3(kBlock), 0, s, nRU, rkout, c->nU, c->d_useg);
        if (int r = check_launch(c, "U pattern")) return r;
        HIPCHK(c, hipStreamSynchronize(s));
        dfree(rkin); dfree(rkout); dfree(rvin);
    }
    // ---- A-dof map (gradients ∂L/∂A of the EulerBeam3DA elements), one entry per element
    if (nRA > 0) {
        HIPCHK(c, hipMalloc(&rkin, sizeof(unsigned int) * nRA));
        HIPCHK(c, hipMalloc(&rkout, sizeof(unsigned int) * nRA));
        HIPCHK(c, hipMalloc(&rvin, sizeof(uint32_t) * nRA));
        HIPCHK(c, hipMalloc(&c->d_asrc, sizeof(uint32_t) * nRA));
        HIPCHK(c, hipMalloc(&c->d_aseg, sizeof(int32_t) * (c->nA + 1)));
        for (auto& g : c->groups) {
            if (!g.amask || g.nele == 0) continue;
            hipLaunchKernelGGL(gen_rpairs_kernel, dim3(grid(g.nele)), dim3(kBlock), 0, s, g.nele, 1, g.d_aidx, g.abase,
                               rkin + g.abase, rvin + g.abase);
        }
        int abit = 1;
        while (abit < 32 && (unsigned long long)c->nA > (1ull << abit)) ++abit;
        tmp_bytes = 0;
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, rkin, rkout, rvin, c->d_asrc, (int)nRA, 0, abit, s));
        HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
        HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, rkin, rkout, rvin, c->d_asrc, (int)nRA, 0, abit, s));
        dfree(tmp);
        hipLaunchKernelGGL(rseg_kernel, dim3(grid(nRA + 1)), dim3(kBlock), 0, s, nRA, rkout, c->nA, c->d_aseg);
        if (int r = check_launch(c, "A pattern")) return r;
        HIPCHK(c, hipStreamSynchronize(s));
        dfree(rkin); dfree(rkout); dfree(rvin);
    }
    // ---- tuner: time each shape of the nz scatter on these maps and keep the fastest.
    // The sums are identical for every shape (each segment is summed in the same order).
    if (nnz > 0) {
        float best = 1e30f;
        for (int v = 0; v < kScatterVariants; ++v) {
            float tmin = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                HIPCHK(c, hipEventRecord(c->ev[0], s));
                if (coded) launch_scatter<true>(v, s, nnz, c->d_kseg, c->d_ksrc, c->d_Kbuf, c->d_nzval);
                else launch_scatter<false>(v, s, nnz, c->d_kseg, c->d_ksrc, c->d_Kbuf, c->d_nzval);
                HIPCHK(c, hipEventRecord(c->ev[1], s));
                HIPCHK(c, hipEventSynchronize(c->ev[1]));
                float ms = 0.f;
                HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                if (rep > 0 && ms < tmin) tmin = ms;
            }
            if (int r = check_launch(c, "scatter tuning")) return r;
            c->scatter_ms[v] = tmin;
            if (tmin < best) { best = tmin; c->scatter_variant = v; }
        }
    }
    c->beam_tuned = false;
    c->prepared = true;
    if (nnz_out) *nnz_out = nnz;
    return MH_OK;
}

int mh_get_pattern(mh_ctx* c, int64_t* colptr, int64_t* rowval) {
    if (!c || !c->prepared) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int32_t> cp(c->nX + 1), rv(c->nnz);
    HIPCHK(c, copy_sync(c, cp.data(), c->d_colptr, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost));
    if (c->nnz) HIPCHK(c, copy_sync(c, rv.data(), c->d_rowval, sizeof(int32_t) * rv.size(), hipMemcpyDeviceToHost));
    if (colptr) for (size_t i = 0; i < cp.size(); ++i) colptr[i] = (int64_t)cp[i] + 1;
    if (rowval) for (size_t i = 0; i < rv.size(); ++i) rowval[i] = (int64_t)rv[i] + 1;
    return MH_OK;
}

int mh_get_asm(mh_ctx* c, int igroup, int64_t* asm1, int64_t* asm2) {
    if (!c || !c->prepared || igroup < 0 || igroup >= (int)c->groups.size()) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const Group& g = c->groups[igroup];
    if (asm1) std::copy(g.h_idx.begin(), g.h_idx.end(), asm1);
    if (asm2 && g.nX > 0 && g.nele > 0) {
        const int64_t per = (int64_t)g.nX * g.nX, n = g.nele * per;
        int32_t* d = nullptr;
        HIPCHK(c, hipMalloc(&d, sizeof(int32_t) * n));
        hipLaunchKernelGGL(asm_dump_kernel, dim3(grid(n)), dim3(kBlock), 0, c->stream, g.nele, g.nX, g.d_idx, c->d_colptr,
                           c->d_rowval, d);
        std::vector<int32_t> h(n);
        HIPCHK(c, hipMemcpyAsync(h.data(), d, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(d);
        for (int64_t q = 0; q < n; ++q) asm2[q] = h[q];   // q = e*nX² + i + nX*j
    }
    return MH_OK;
}

int mh_set_X(mh_ctx* c, int ider, const double* x) {
    if (!c || ider < 0 || ider > 2 || !x) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nX) HIPCHK(c, hipMemcpyAsync(c->d_X[ider], x, sizeof(double) * c->nX, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}
int mh_set_U(mh_ctx* c, const double* u) {
    if (!c || (c->nU && !u)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nU) HIPCHK(c, hipMemcpyAsync(c->d_U, u, sizeof(double) * c->nU, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}
int mh_set_A(mh_ctx* c, const double* a) {
    if (!c || (c->nA && !a)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nA) HIPCHK(c, hipMemcpyAsync(c->d_A, a, sizeof(double) * c->nA, hipMemcpyHostToDevice, c->stream));
    if (int r = apply_ascale(c)) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}
int mh_set_point_values(mh_ctx* c, int igroup, const double* v) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size() || !v) return MH_ERR_ARG;
    Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_DOFLOAD && g.kind != MH_KIND_HOLD) { c->err = "group has no host-evaluated values"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    if (g.nele) HIPCHK(c, hipMemcpyAsync(g.d_vals, v, sizeof(double) * g.nele, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}

int mh_set_time(mh_ctx* c, double t) {
    if (!c) return MH_ERR_ARG;
    c->t = t;
    return MH_OK;
}

}  // extern "C"

// Every launch of one assemble!{mission} of AssemblySweepX{OX}: the NaN-flag reset, the
// element kernels of all groups, the nz scatter and the rhs scatter, on the context's
// stream. With timed, HIP events bracket the three phases (not inside a graph capture).
// The tuner of the static beam kernel's LDS tile (elements per workgroup), run once per
// prepared context at its first assembly (the device state is then the caller's), before any
// graph capture. Every shape gives the same numbers.
static int tune_beams(Ctx* c) {
    c->beam_tuned = true;
    c->beam_variant = 1;
    bool any_static = false;
    for (auto& g : c->groups) any_static = any_static || (g.compact && g.nele > 0);
    if (!any_static) return MH_OK;
    hipStream_t s = c->stream;
    float best = 1e30f;
    for (int v = 0; v < kBeamVariants; ++v) {
        float tmin = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            HIPCHK(c, hipEventRecord(c->ev[0], s));
            for (size_t gi = 0; gi < c->groups.size(); ++gi) {
                const Group& g = c->groups[gi];
                if (!g.compact || g.nele == 0) continue;
                BeamArgs a = static_beam_args(c, g, (unsigned long long)gi << 40);
                launch_beam_static(v, a, g.kind == MH_KIND_BEAM_U, s);
            }
            HIPCHK(c, hipEventRecord(c->ev[1], s));
            HIPCHK(c, hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            if (rep > 0 && ms < tmin) tmin = ms;
        }
        if (int r = check_launch(c, "beam kernel tuning")) return r;
        c->beam_ms[v] = tmin;
        if (tmin < best) { best = tmin; c->beam_variant = v; }
    }
    return MH_OK;
}

// fused: Lλx goes straight into the chain solve's blocks (each nz to its slot; the pc part
// cleared first) instead of nzval, for a device Newton iteration that solves right after.
static int enqueue_sweepx(Ctx* c, int OX, bool step, const double* c6, bool timed, bool fused = false) {
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, s, c->d_flag, ~0ull);
    if (timed) HIPCHK(c, hipEventRecord(c->ev[0], s));
    PointBatch pb;
    pb.ngroup = 0;
    pb.start[0] = 0;
    auto flush_points = [&]() {
        if (pb.ngroup == 0) return;
        hipLaunchKernelGGL(point_kernel, dim3(grid(pb.start[pb.ngroup])), dim3(kBlock), 0, s, pb);
        pb.ngroup = 0;
    };
    for (size_t gi = 0; gi < c->groups.size(); ++gi) {
        Group& g = c->groups[gi];
        if (g.nX == 0 || g.nele == 0) continue;
        const unsigned long long gtag = (unsigned long long)gi << 40;
        if (g.kind == MH_KIND_BAR || g.kind == MH_KIND_BAR_U) {
            if (int r = launch_bars(c, g, OX, step, c6, gtag)) return r;
        } else if ((g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U) && OX > 0) {
            if (int r = launch_dyn_beams(c, g, OX, step, c6, gtag)) return r;
        } else if (g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U) {
            BeamArgs a = static_beam_args(c, g, gtag);
            launch_beam_static(c->beam_variant, a, g.kind == MH_KIND_BEAM_U, s);
        } else {
            PointArgs& a = pb.g[pb.ngroup];
            a.kind = g.kind; a.nele = g.nele; a.nX = g.nX; a.idx = g.d_idx; a.X = c->d_X[0]; a.vals = g.d_vals;
            for (int k = 0; k < 12; ++k) a.s[k] = g.scaleX[k];
            a.K = c->d_Kbuf + g.kbase; a.R = c->d_Rbuf + g.rbase; a.dR = step ? c->d_dRbuf + g.rbase : nullptr;
            a.flag = c->d_flag; a.gtag = gtag;
            pb.start[pb.ngroup + 1] = pb.start[pb.ngroup] + g.nele;
            if (++pb.ngroup == kPointBatch) flush_points();
        }
    }
    flush_points();
    if (int r = check_launch(c, "element kernels")) return r;
    if (timed) HIPCHK(c, hipEventRecord(c->ev[1], s));
    if (c->nnz) {
        double* out = c->d_nzval;
        const int32_t* dst = nullptr;
        if (fused) {
            const int64_t pcbase = 3 * c->cnb * c->cB * c->cB + c->cnb * c->cB;
            if (c->cnp)
                hipLaunchKernelGGL(chain::fill_f64, dim3((unsigned)((4 * c->cnp + chain::kCB - 1) / chain::kCB)),
                                   dim3(chain::kCB), 0, s, 4 * c->cnp, c->d_cbuf + pcbase, 0.0);
            out = c->d_cbuf;
            dst = c->d_cslot;
        }
        if (c->coded) launch_scatter<true>(c->scatter_variant, s, c->nnz, c->d_kseg, c->d_ksrc, c->d_Kbuf, out, dst);
        else launch_scatter<false>(c->scatter_variant, s, c->nnz, c->d_kseg, c->d_ksrc, c->d_Kbuf, out, dst);
    }
    if (timed) HIPCHK(c, hipEventRecord(c->ev[2], s));
    if (c->nX) {
        const int64_t per = ((c->nX + kTile - 1) / kTile + kXCD - 1) / kXCD;
        if (step)
            hipLaunchKernelGGL((scatter_kernel<true, false>), dim3((unsigned)(per * kXCD)), dim3(kTile), 0, s, c->nX, per,
                               c->d_rseg, c->d_rsrc, c->d_Rbuf, (const double*)c->d_dRbuf, c->d_rhs);
        else
            hipLaunchKernelGGL((scatter_kernel<false, false>), dim3((unsigned)(per * kXCD)), dim3(kTile), 0, s, c->nX, per,
                               c->d_rseg, c->d_rsrc, c->d_Rbuf, (const double*)nullptr, c->d_rhs);
    }
    if (timed) HIPCHK(c, hipEventRecord(c->ev[3], s));
    // the NaN flag goes to pinned host memory on the stream (a node of a captured graph)
    HIPCHK(c, hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    return check_launch(c, "scatter kernels");
}

// Whether a kernel of the assembly takes the time as an argument (bars: their loads may vary
// in time); beams, Holds and DofLoads read it only through values set by the host.
static bool reads_time(const Ctx* c) {
    for (const auto& g : c->groups)
        if ((g.kind == MH_KIND_BAR || g.kind == MH_KIND_BAR_U) && g.nele) return true;
    return false;
}

// Wait for the assembly, then report the first element that returned NaN.
static int finish_sweepx(Ctx* c, mh_err* err) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const unsigned long long flag = *c->h_flag;
    if (flag != ~0ull) {
        int gi = (int)(flag >> 40);
        int64_t e = (int64_t)(flag & ((1ull << 40) - 1));
        c->err = "element " + std::to_string(e + 1) + " of element type " + std::to_string(gi + 1) + " returned NaN";
        if (err) { err->code = MH_ERR_NAN; err->igroup = gi; err->iele = e; }
        return MH_ERR_NAN;
    }
    return MH_OK;
}

// The Newmark coefficients of this call into d_c6, ahead of the kernels on the stream: kernels
// (eager or replayed from a graph) read them there, so a graph outlives a change of Δt.
static int upload_c6(Ctx* c, int OX, const double* c6) {
    for (int i = 0; i < 6; ++i) c->h_c6[i] = (OX > 0 && c6) ? c6[i] : 0.;
    HIPCHK(c, hipMemcpyAsync(c->d_c6, c->h_c6, sizeof(double) * 6, hipMemcpyHostToDevice, c->stream));
    return MH_OK;
}

static int check_sweepx_args(Ctx* c, int OX, int mission, const double* c6, mh_err* err) {
    if (err) { err->code = MH_OK; err->igroup = -1; err->iele = -1; }
    if (!c || !c->prepared || OX < 0 || OX > 2 || mission < 0 || mission > 1 || (OX > 0 && !c6)) return MH_ERR_ARG;
    if (OX != c->OX) {
        c->err = "assembler prepared for OX=" + std::to_string(c->OX) + ", called with OX=" + std::to_string(OX);
        return MH_ERR_ARG;
    }
    return MH_OK;
}

extern "C" {

int mh_assemble_sweepx(mh_ctx* c, int OX, int mission, const double* c6, mh_err* err) {
    if (int r = check_sweepx_args(c, OX, mission, c6, err)) return r;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->beam_tuned)
        if (int r = tune_beams(c)) return r;
    const bool step = (mission == 0) && OX > 0;
    if (int r = upload_c6(c, OX, c6)) return r;
    if (int r = enqueue_sweepx(c, OX, step, c6, true)) return r;
    c->timed = true;
    c->graph_timed = false;
    return finish_sweepx(c, err);
}

int mh_assemble_sweepx_graph(mh_ctx* c, int OX, int mission, const double* c6, mh_err* err) {
    if (int r = check_sweepx_args(c, OX, mission, c6, err)) return r;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->beam_tuned) {   // before any capture: the tuner times launches
        if (int r = tune_beams(c)) return r;
        ++c->gen;
    }
    const bool step = (mission == 0) && OX > 0;
    if (int r = upload_c6(c, OX, c6)) return r;
    const double key[9] = {0, 0, 0, 0, 0, 0, 0, 0, reads_time(c) ? c->t : 0.};
    const int sl = step ? 1 : 0;
    bool same = c->gexec[sl] && c->gOX[sl] == OX && c->ggen[sl] == c->gen;
    for (int i = 0; i < 9 && same; ++i) same = std::memcmp(&key[i], &c->gkey[sl][i], sizeof(double)) == 0;
    hipStream_t s = c->stream;
    if (!same) {
        free_graph(c, sl);
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int r = enqueue_sweepx(c, OX, step, c6, false);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (r != MH_OK || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (r == MH_OK) c->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec);
            return r != MH_OK ? r : MH_ERR_HIP;
        }
        c->graph[sl] = g;
        HIPCHK(c, hipGraphInstantiateWithFlags(&c->gexec[sl], c->graph[sl], 0));
        std::memcpy(c->gkey[sl], key, sizeof(key));
        c->gOX[sl] = OX; c->ggen[sl] = c->gen;
        ++c->graph_captures;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    HIPCHK(c, hipGraphLaunch(c->gexec[sl], s));
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    c->timed = true;
    c->graph_timed = true;
    return finish_sweepx(c, err);
}

int mh_graph_captures(mh_ctx* c, int64_t* n) {
    if (!c || !n) return MH_ERR_ARG;
    *n = c->graph_captures;
    return MH_OK;
}

int mh_set_lambda(mh_ctx* c, const double* lam) {
    if (!c || !lam || !(c->prepared || c->direct)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->d_lam) HIPCHK(c, hipMalloc(&c->d_lam, sizeof(double) * std::max<int64_t>(c->nX, 1)));
    if (c->nX) HIPCHK(c, hipMemcpyAsync(c->d_lam, lam, sizeof(double) * c->nX, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}

int mh_vjp_lagrangian(mh_ctx* c, int OX, mh_err* err) {
    if (err) { err->code = MH_OK; err->igroup = -1; err->iele = -1; }
    if (!c || !c->prepared || !c->d_lam || OX < 0 || OX > 2) return MH_ERR_ARG;
    for (auto& g : c->groups)
        if (g.kind == MH_KIND_HOST && g.nele) { c->err = "the adjoint does not cover host-evaluated elements"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int nd = OX + 1;
    if (!c->d_gbuf) {
        HIPCHK(c, hipMalloc(&c->d_gbuf, sizeof(double) * std::max<int64_t>(3 * c->nR, 1)));
        HIPCHK(c, hipMalloc(&c->d_gX, sizeof(double) * std::max<int64_t>(3 * c->nX, 1)));
        HIPCHK(c, hipMalloc(&c->d_gubuf, sizeof(double) * std::max<int64_t>(c->nRU, 1)));
        HIPCHK(c, hipMalloc(&c->d_gU, sizeof(double) * std::max<int64_t>(c->nU, 1)));
        if (c->nRA) {
            HIPCHK(c, hipMalloc(&c->d_gabuf, sizeof(double) * c->nRA));
            HIPCHK(c, hipMalloc(&c->d_gA, sizeof(double) * std::max<int64_t>(c->nA, 1)));
        }
    }
    if (OX > 0 && !c->d_Kc) {
        int64_t nKc = 0;
        for (auto& g : c->groups)
            if (g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U) { g.kcbase = nKc; nKc += g.nele * mh::kCompact; }
        if (nKc) HIPCHK(c, hipMalloc(&c->d_Kc, sizeof(double) * nKc));
        if (nKc && !c->d_Kb) HIPCHK(c, hipMalloc(&c->d_Kb, sizeof(double) * (nKc / mh::kCompact * mh::kDynBase)));
    }
    if (OX > 0 && !c->d_ybuf) {
        int64_t nb = 0;   // beams, in the layout of Kc
        for (auto& g : c->groups)
            if (g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U) nb = std::max(nb, g.kcbase / mh::kCompact + g.nele);
        if (nb) HIPCHK(c, hipMalloc(&c->d_ybuf, sizeof(double) * 15 * nb));
    }
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, s, c->d_flag, ~0ull);
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    for (size_t gi = 0; gi < c->groups.size(); ++gi) {
        const Group& g = c->groups[gi];
        if (g.nX == 0 || g.nele == 0) continue;
        if (int r = launch_vjp(c, g, nd, (unsigned long long)gi << 40)) return r;
    }
    HIPCHK(c, hipEventRecord(c->ev[1], s));
    if (c->nX) {
        const int64_t per = ((c->nX + kTile - 1) / kTile + kXCD - 1) / kXCD;
        for (int d = 0; d < nd; ++d)
            hipLaunchKernelGGL((scatter_kernel<false, false>), dim3((unsigned)(per * kXCD)), dim3(kTile), 0, s, c->nX, per,
                               c->d_rseg, c->d_rsrc, c->d_gbuf + d * c->nR, (const double*)nullptr, c->d_gX + d * c->nX);
    }
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    if (c->nU) {
        if (c->nRU > 16 * c->nU) {   // long segments: a workgroup per U-dof
            hipLaunchKernelGGL(scatter_long_kernel, dim3((unsigned)c->nU), dim3(kTile), 0, s, c->nU, c->d_useg, c->d_usrc,
                               c->d_gubuf, c->d_gU);
        } else if (c->nRU) {
            const int64_t per = ((c->nU + kTile - 1) / kTile + kXCD - 1) / kXCD;
            hipLaunchKernelGGL((scatter_kernel<false, false>), dim3((unsigned)(per * kXCD)), dim3(kTile), 0, s, c->nU, per,
                               c->d_useg, c->d_usrc, c->d_gubuf, (const double*)nullptr, c->d_gU);
        } else {
            HIPCHK(c, hipMemsetAsync(c->d_gU, 0, sizeof(double) * c->nU, s));
        }
    }
    if (c->nRA) {   // ∂L/∂A: few A-dofs shared by many elements, a workgroup per A-dof
        hipLaunchKernelGGL(scatter_long_kernel, dim3((unsigned)c->nA), dim3(kTile), 0, s, c->nA, c->d_aseg, c->d_asrc,
                           c->d_gabuf, c->d_gA);
    }
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    if (int r = check_launch(c, "adjoint scatter")) return r;
    unsigned long long flag = 0;
    HIPCHK(c, hipMemcpyAsync(&flag, c->d_flag, sizeof(flag), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->timed = true;
    c->vjp_nd = nd;
    if (flag != ~0ull) {
        int gi = (int)(flag >> 40);
        int64_t e = (int64_t)(flag & ((1ull << 40) - 1));
        c->err = "element " + std::to_string(e + 1) + " of element type " + std::to_string(gi + 1) + " returned NaN";
        if (err) { err->code = MH_ERR_NAN; err->igroup = gi; err->iele = e; }
        return MH_ERR_NAN;
    }
    return MH_OK;
}

int mh_get_vjp(mh_ctx* c, double* gX, double* gU, double* gA) {
    if (!c || !c->prepared || c->vjp_nd == 0) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (gX && c->nX)
        HIPCHK(c, hipMemcpyAsync(gX, c->d_gX, sizeof(double) * c->nX * c->vjp_nd, hipMemcpyDeviceToHost, c->stream));
    if (gU && c->nU) HIPCHK(c, hipMemcpyAsync(gU, c->d_gU, sizeof(double) * c->nU, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (gA && c->nA && c->nRA) {
        HIPCHK(c, hipMemcpyAsync(gA, c->d_gA, sizeof(double) * c->nA, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else if (gA) {
        for (int64_t i = 0; i < c->nA; ++i) gA[i] = 0.;   // no element here depends on A
    }
    return MH_OK;
}

int mh_get_rhs(mh_ctx* c, double* rhs) {
    if (!c || !c->prepared || !rhs) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nX) HIPCHK(c, hipMemcpyAsync(rhs, c->d_rhs, sizeof(double) * c->nX, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}
int mh_get_nzval(mh_ctx* c, double* nzval) {
    if (!c || !c->prepared || !nzval) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nnz) HIPCHK(c, hipMemcpyAsync(nzval, c->d_nzval, sizeof(double) * c->nnz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}
int mh_device_outputs(mh_ctx* c, double** d_rhs, double** d_nzval) {
    if (!c || !c->prepared) return MH_ERR_ARG;
    if (d_rhs) *d_rhs = c->d_rhs;
    if (d_nzval) *d_nzval = c->d_nzval;
    return MH_OK;
}
int mh_get_element_blocks(mh_ctx* c, int igroup, double* R, double* K) {
    if (!c || !c->prepared || igroup < 0 || igroup >= (int)c->groups.size()) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const Group& g = c->groups[igroup];
    const int64_t n = g.nele, nX = g.nX;
    if (n == 0 || nX == 0) return MH_OK;
    std::vector<double> hk(n * g.kper), hr(n * nX);
    HIPCHK(c, hipMemcpyAsync(hk.data(), c->d_Kbuf + g.kbase, sizeof(double) * hk.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hr.data(), c->d_Rbuf + g.rbase, sizeof(double) * hr.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t e = 0; e < n; ++e) {
        if (R) for (int64_t i = 0; i < nX; ++i) R[e * nX + i] = hr[i * n + e];
        if (K)
            for (int64_t j = 0; j < nX; ++j)
                for (int64_t i = 0; i < nX; ++i) {
                    double v;
                    if (g.compact) { double m; int slot = mh::compact_slot((int)i, (int)j, &m); v = m * hk[e * g.kper + slot]; }
                    else v = hk[e * g.kper + i + nX * j];
                    K[e * nX * nX + i + nX * j] = v;
                }
    }
    return MH_OK;
}

int mh_last_timings(mh_ctx* c, double* ms) {
    if (!c || !ms || !c->timed) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    float t = 0.f;
    if (c->graph_timed) {   // one graph launch: its whole duration, phases not separated
        HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
        ms[0] = t; ms[1] = 0.; ms[2] = 0.;
        return MH_OK;
    }
    for (int i = 0; i < 3; ++i) {
        HIPCHK(c, hipEventElapsedTime(&t, c->ev[i], c->ev[i + 1]));
        ms[i] = t;
    }
    return MH_OK;
}

int mh_calib_copy(mh_ctx* c, int64_t n, double* ms) {
    if (!c || n <= 0) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    double *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, sizeof(double) * n));
    HIPCHK(c, hipMalloc(&b, sizeof(double) * n));
    HIPCHK(c, hipMemsetAsync(a, 0, sizeof(double) * n, c->stream));
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(kBlock), 0, c->stream, n, a, b);   // warm-up
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(kBlock), 0, c->stream, n, a, b);
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    if (ms) *ms = t;
    dfree(a); dfree(b);
    return MH_OK;
}

int mh_set_scatter_variant(mh_ctx* c, int variant) {
    if (!c || !c->prepared || variant < 0 || variant >= kScatterVariants) return MH_ERR_ARG;
    c->scatter_variant = variant;
    ++c->gen;   // a captured graph holds the old launch
    return MH_OK;
}

int mh_get_tuning(mh_ctx* c, int* variant, float* ms) {
    if (!c || !c->prepared) return MH_ERR_ARG;
    if (variant) *variant = c->scatter_variant;
    if (ms) for (int v = 0; v < kScatterVariants; ++v) ms[v] = c->scatter_ms[v];
    return MH_OK;
}

int mh_get_beam_tuning(mh_ctx* c, int* block, float* ms) {
    if (!c || !c->prepared || !c->beam_tuned) return MH_ERR_ARG;
    if (block) *block = kBeamVariantBlock[c->beam_variant];
    if (ms) for (int v = 0; v < kBeamVariants; ++v) ms[v] = c->beam_ms[v];
    return MH_OK;
}

int mh_sync(mh_ctx* c) {
    if (!c) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}


// ============================================================================ AssemblyDirect
int mh_set_group_scales(mh_ctx* c, int igroup, const double* scaleL, const double* scaleU) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size()) return MH_ERR_ARG;
    Group& g = c->groups[igroup];
    if (scaleL) for (int k = 0; k < g.nX && k < 12; ++k) g.scaleL[k] = scaleL[k];
    if (scaleU) for (int k = 0; k < g.nU && k < 3; ++k) g.scaleU[k] = scaleU[k];
    ++c->gen;
    return MH_OK;
}

int mh_prepare_direct(mh_ctx* c, int OX, int OU, int IA, int64_t* nnz4) {
    if (!c || OX < 0 || OX > 2 || OU < 0 || OU > 2 || IA < 0 || IA > 1) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    free_prepared(c);
    // A-dofs reach the device only through A-scaled beams: their section follows A (mh_set_A)
    // and their A column comes from mh_beam_dRdA, outside these blocks
    int64_t nR = 0, nKd = 0, nKU = 0, nKc = 0;
    for (auto& g : c->groups) {
        if (g.kind == MH_KIND_ZERO || g.kind == MH_KIND_HOST) {
            c->err = "elements with user functions (DofCost, DofConstraint) are not evaluated on the device in AssemblyDirect";
            return MH_ERR_ARG;
        }
        g.rbase = nR; g.kdbase = nKd; g.kubase = nKU; g.kcbase = nKc;
        nR += g.nele * g.nX;
        nKd += g.nele * g.nX * g.nX;
        nKU += g.nele * g.nX * g.nU;
        if (g.kind == MH_KIND_BEAM || g.kind == MH_KIND_BEAM_U) nKc += g.nele * mh::kCompact;
    }
    if (nR >= (int64_t)UINT32_MAX || nKd >= (int64_t)UINT32_MAX) { c->err = "too many element outputs for 32-bit maps"; return MH_ERR_SIZE; }
    c->nR = nR; c->nKd = nKd; c->nKU = nKU;
    c->dOX = OX; c->dOU = OU; c->dIA = IA;
    if (int r = build_rmap(c)) return r;
    std::vector<PairSpec> xx, xu, ux, uu;
    for (auto& g : c->groups) {
        if (g.nX == 0 || g.nele == 0) continue;
        xx.push_back({g.nele, g.nX, g.d_idx, g.nX, g.d_idx, g.kdbase});
        if (g.nU) {
            xu.push_back({g.nele, g.nX, g.d_idx, g.nU, g.d_uidx, g.kubase});
            ux.push_back({g.nele, g.nU, g.d_uidx, g.nX, g.d_idx, g.kubase});
            uu.push_back({g.nele, g.nU, g.d_uidx, g.nU, g.d_uidx, 0});
        }
    }
    if (int r = build_dmap(c, xx, c->nX, c->nX, true, true, c->dXX)) return r;
    if (int r = build_dmap(c, xu, c->nX, c->nU, true, false, c->dXU)) return r;
    if (int r = build_dmap(c, ux, c->nU, c->nX, false, true, c->dUX)) return r;
    if (int r = build_dmap(c, uu, c->nU, c->nU, false, false, c->dUU)) return r;
    HIPCHK(c, hipMalloc(&c->d_Kd, sizeof(double) * std::max<int64_t>((OX + 1) * nKd, 1)));
    HIPCHK(c, hipMalloc(&c->d_KU, sizeof(double) * std::max<int64_t>(nKU, 1)));
    HIPCHK(c, hipMalloc(&c->d_Kc, sizeof(double) * std::max<int64_t>(nKc, 1)));
    HIPCHK(c, hipMalloc(&c->d_Rd, sizeof(double) * std::max<int64_t>(nR, 1)));
    HIPCHK(c, hipMalloc(&c->d_L1x, sizeof(double) * std::max<int64_t>(nR, 1)));
    // residual-path elements (beams, bars) add nothing to L1[X] and never write their slots
    HIPCHK(c, hipMemsetAsync(c->d_L1x, 0, sizeof(double) * std::max<int64_t>(nR, 1), c->stream));
    HIPCHK(c, hipMalloc(&c->d_dL1, sizeof(double) * std::max<int64_t>(2 * c->nX, 1)));
    HIPCHK(c, hipMalloc(&c->d_dL2, sizeof(double) * std::max<int64_t>(2 * (OX + 1) * c->dXX.nnz + 2 * c->dXU.nnz, 1)));
    c->direct = true;
    if (nnz4) { nnz4[0] = c->dXX.nnz; nnz4[1] = c->dXU.nnz; nnz4[2] = c->dUX.nnz; nnz4[3] = c->dUU.nnz; }
    return MH_OK;
}

int mh_get_direct_pattern(mh_ctx* c, int which, int64_t* colptr, int64_t* rowval) {
    if (!c || !c->direct || which < 0 || which > 3) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const DMap& M = which == 0 ? c->dXX : which == 1 ? c->dXU : which == 2 ? c->dUX : c->dUU;
    std::vector<int32_t> cp(M.ncol + 1, 0), rv(M.nnz);
    if (M.colptr) HIPCHK(c, copy_sync(c, cp.data(), M.colptr, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost));
    if (M.nnz) HIPCHK(c, copy_sync(c, rv.data(), M.rowval, sizeof(int32_t) * rv.size(), hipMemcpyDeviceToHost));
    if (colptr) for (size_t i = 0; i < cp.size(); ++i) colptr[i] = (int64_t)cp[i] + 1;
    if (rowval) for (size_t i = 0; i < rv.size(); ++i) rowval[i] = (int64_t)rv[i] + 1;
    return MH_OK;
}

int mh_assemble_direct(mh_ctx* c, int mission, double dt, mh_err* err) {
    if (err) { err->code = MH_OK; err->igroup = -1; err->iele = -1; }
    if (!c || !c->direct || mission < 0 || mission > 1) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& g : c->groups)
        if (g.kind == MH_KIND_HOLD && g.nele && !c->d_lam) { c->err = "AssemblyDirect needs Λ (mh_set_lambda)"; return MH_ERR_ARG; }
    hipStream_t s = c->stream;
    const int nd = c->dOX + 1;
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, s, c->d_flag, ~0ull);
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    for (size_t gi = 0; gi < c->groups.size(); ++gi) {
        const Group& g = c->groups[gi];
        if (g.nX == 0 || g.nele == 0) continue;
        if (int r = launch_direct(c, g, nd, dt, (unsigned long long)gi << 40)) return r;
    }
    HIPCHK(c, hipEventRecord(c->ev[1], s));
    const double* nul = nullptr;
    if (c->nX) {
        const int64_t per = ((c->nX + kTile - 1) / kTile + kXCD - 1) / kXCD;
        hipLaunchKernelGGL((scatter_kernel<false, false>), dim3((unsigned)(per * kXCD)), dim3(kTile), 0, s, c->nX, per,
                           c->d_rseg, c->d_rsrc, c->d_Rd, nul, c->d_dL1);
        hipLaunchKernelGGL((scatter_kernel<false, false>), dim3((unsigned)(per * kXCD)), dim3(kTile), 0, s, c->nX, per,
                           c->d_rseg, c->d_rsrc, c->d_L1x, nul, c->d_dL1 + c->nX);
    }
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    if (mission == 1) {
        const int64_t nxx = c->dXX.nnz, nxu = c->dXU.nnz;
        for (int d = 0; d < nd; ++d)
            if (nxx) {
                launch_scatter<false>(c->scatter_variant, s, nxx, c->dXX.seg, c->dXX.src, c->d_Kd + d * c->nKd,
                                      c->d_dL2 + d * nxx);
                launch_scatter<false>(c->scatter_variant, s, nxx, c->dXX.seg, c->dXX.srcT, c->d_Kd + d * c->nKd,
                                      c->d_dL2 + (nd + d) * nxx);
            }
        if (nxu) {
            launch_scatter<false>(c->scatter_variant, s, nxu, c->dXU.seg, c->dXU.src, c->d_KU, c->d_dL2 + 2 * nd * nxx);
            launch_scatter<false>(c->scatter_variant, s, c->dUX.nnz, c->dUX.seg, c->dUX.srcT, c->d_KU,
                                  c->d_dL2 + 2 * nd * nxx + nxu);
        }
    }
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    if (int r = check_launch(c, "direct scatter")) return r;
    unsigned long long flag = 0;
    HIPCHK(c, hipMemcpyAsync(&flag, c->d_flag, sizeof(flag), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    c->timed = true;
    if (flag != ~0ull) {
        int gi = (int)(flag >> 40);
        int64_t e = (int64_t)(flag & ((1ull << 40) - 1));
        c->err = "element " + std::to_string(e + 1) + " of element type " + std::to_string(gi + 1) + " returned NaN";
        if (err) { err->code = MH_ERR_NAN; err->igroup = gi; err->iele = e; }
        return MH_ERR_NAN;
    }
    return MH_OK;
}

int mh_get_direct(mh_ctx* c, double* L1, double* L2) {
    if (!c || !c->direct) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t nX = c->nX, nU = c->nU;
    const int ndX = c->dOX + 1, ndU = c->dOU + 1;
    const int64_t nxx = c->dXX.nnz, nxu = c->dXU.nnz, nux = c->dUX.nnz, nuu = c->dUU.nnz;
    if (L1) {
        const int64_t n1 = nX + (int64_t)ndX * nX + (int64_t)ndU * nU;
        std::fill(L1, L1 + n1, 0.);
        if (nX) {
            HIPCHK(c, copy_sync(c, L1, c->d_dL1, sizeof(double) * nX, hipMemcpyDeviceToHost));
            HIPCHK(c, copy_sync(c, L1 + nX, c->d_dL1 + nX, sizeof(double) * nX, hipMemcpyDeviceToHost));
        }
    }
    if (L2) {
        // pairs (α,β) ∈ {Λ,X,U}² row-major, (Λ,Λ) has no matrices; nder(α)×nder(β) blocks each
        const int nder[3] = {1, ndX, ndU};
        const int64_t nnzp[3][3] = {{0, nxx, nxu}, {nxx, nxx, nxu}, {nux, nux, nuu}};
        int64_t off = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                if (a == 0 && b == 0) continue;
                const int64_t blk = nnzp[a][b];
                for (int i = 0; i < nder[a]; ++i)
                    for (int j = 0; j < nder[b]; ++j) {
                        double* dst = L2 + off + ((int64_t)i * nder[b] + j) * blk;
                        const double* src = nullptr;
                        if (a == 0 && b == 1 && i == 0) src = c->d_dL2 + j * nxx;
                        if (a == 1 && b == 0 && j == 0) src = c->d_dL2 + (ndX + i) * nxx;
                        if (a == 0 && b == 2 && i == 0 && j == 0) src = c->d_dL2 + 2 * ndX * nxx;
                        if (a == 2 && b == 0 && i == 0 && j == 0) src = c->d_dL2 + 2 * ndX * nxx + nxu;
                        if (src && blk) HIPCHK(c, copy_sync(c, dst, src, sizeof(double) * blk, hipMemcpyDeviceToHost));
                        else std::fill(dst, dst + blk, 0.);
                    }
                off += (int64_t)nder[a] * nder[b] * blk;
            }
    }
    return MH_OK;
}
}  // extern "C"

// ------------------------------------------------------------------ chain solve
namespace {

int64_t cgrid(int64_t n) { return (n + chain::kCB - 1) / chain::kCB; }

// fused: the assembly already scattered Lλx into the blocks (enqueue_sweepx with fused),
// after clearing pc; the pack is skipped.
template <int B>
int enqueue_chain_solve(Ctx* c, const double* rhs, bool timed = true, bool fused = false) {
    using namespace chain;
    hipStream_t s = c->stream;
    const int64_t nb = c->cnb, BB = (int64_t)B * B;
    double* D = c->d_cbuf;
    double* L = D + nb * BB;
    double* U = L + nb * BB;
    double* b = U + nb * BB;
    double* pc = b + nb * B;
    if (timed) HIPCHK(c, hipEventRecord(c->cev[0], s));
    hipLaunchKernelGGL(clear_flag, dim3(1), dim3(1), 0, s, c->d_cflag);
    if (c->cnp && !fused) hipLaunchKernelGGL(fill_f64, dim3(cgrid(4 * c->cnp)), dim3(kCB), 0, s, 4 * c->cnp, pc, 0.0);
    if (c->cnu) hipLaunchKernelGGL(zero_slots_kernel, dim3(cgrid(c->cnu)), dim3(kCB), 0, s, c->cnu, c->d_cunc, D);
    if (c->cnd) hipLaunchKernelGGL(dummy_kernel<B>, dim3(cgrid(c->cnd)), dim3(kCB), 0, s, c->cnd, c->d_cdum, D, b);
    if (c->nnz && !fused)
        hipLaunchKernelGGL(pack_kernel, dim3(cgrid(c->nnz)), dim3(kCB), 0, s, c->nnz, c->d_cslot, c->d_nzval, c->d_cbuf);
    hipLaunchKernelGGL(pack_rhs_kernel, dim3(cgrid(c->nX)), dim3(kCB), 0, s, c->nX, c->d_cpos, rhs, b, pc);
    if (c->cnp) hipLaunchKernelGGL(pair_kernel<B>, dim3(cgrid(c->cnp)), dim3(kCB), 0, s, c->cnp, c->d_chpos, pc, D, L,
                                   U, b, c->d_csaved, c->d_cflag);
    if (timed) HIPCHK(c, hipEventRecord(c->cev[1], s));
    int64_t top = 1;
    for (int64_t st = 1; st < nb; st *= 2) {
        const int64_t nel = (nb - st + 2 * st - 1) / (2 * st), nkeep = (nb + 2 * st - 1) / (2 * st);
        hipLaunchKernelGGL(factor_kernel<B>, dim3(cgrid(nel)), dim3(kCB), 0, s, nb, st, D, L, U, b, c->d_cflag);
        hipLaunchKernelGGL(update_kernel<B>, dim3(cgrid(nkeep * B)), dim3(kCB), 0, s, nb, st, D, L, U, b);
        top = 2 * st;
    }
    hipLaunchKernelGGL(top_kernel<B>, dim3(1), dim3(64), 0, s, D, b, c->d_cflag);
    if (timed) HIPCHK(c, hipEventRecord(c->cev[2], s));
    for (int64_t st = top / 2; st >= 1; st /= 2) {
        const int64_t nel = (nb - st + 2 * st - 1) / (2 * st);
        hipLaunchKernelGGL(back_kernel<B>, dim3(cgrid(nel * B)), dim3(kCB), 0, s, nb, st, L, U, b);
    }
    hipLaunchKernelGGL(unpack_kernel<B>, dim3(cgrid(c->nX)), dim3(kCB), 0, s, c->nX, nb, c->d_cpos, c->d_chpos, pc,
                       c->d_csaved, b, c->d_dx);
    if (timed) HIPCHK(c, hipEventRecord(c->cev[3], s));
    HIPCHK(c, hipGetLastError());
    return MH_OK;
}

// Prepare-time maps: the slot of every nz, and the list of uncovered slots.
template <int B>
int build_chain_slots(Ctx* c) {
    using namespace chain;
    hipStream_t s = c->stream;
    const int64_t nslot = 3 * c->cnb * B * B;
    int32_t *unc = nullptr, *at = nullptr;
    HIPCHK(c, hipMalloc(&c->d_cslot, sizeof(int32_t) * std::max<int64_t>(c->nnz, 1)));
    HIPCHK(c, hipMalloc(&unc, sizeof(int32_t) * (nslot + 1)));
    HIPCHK(c, hipMalloc(&at, sizeof(int32_t) * (nslot + 1)));
    hipLaunchKernelGGL(clear_flag, dim3(1), dim3(1), 0, s, c->d_cflag);
    hipLaunchKernelGGL(fill_i32, dim3(cgrid(nslot + 1)), dim3(kCB), 0, s, nslot + 1, unc, 1);
    hipLaunchKernelGGL(slot_kernel<B>, dim3(cgrid(c->nX)), dim3(kCB), 0, s, c->nX, c->cnb, c->d_colptr, c->d_rowval,
                       c->d_cpos, c->d_chpos, c->d_cslot, unc, c->d_cflag);
    if (c->nnz)
        hipLaunchKernelGGL(absolute_slot_kernel, dim3(cgrid(c->nnz)), dim3(kCB), 0, s, c->nnz,
                           nslot + c->cnb * B, c->d_cslot);
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, unc, at, (int)(nslot + 1), s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, unc, at, (int)(nslot + 1), s));
    int32_t nu = 0;
    HIPCHK(c, hipMemcpyAsync(&nu, at + nslot, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(c->h_cflag, c->d_cflag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(tmp);
    c->cnu = nu;
    HIPCHK(c, hipMalloc(&c->d_cunc, sizeof(int32_t) * std::max<int64_t>(nu, 1)));
    hipLaunchKernelGGL(compact_kernel, dim3(cgrid(nslot)), dim3(kCB), 0, s, nslot, unc, at, c->d_cunc);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(unc); dfree(at);
    if (*c->h_cflag & kBadPattern) { c->err = "matrix entry outside the chain's block-tridiagonal structure"; return MH_ERR_ARG; }
    return MH_OK;
}

}  // namespace

extern "C" {

int mh_chain_prepare(mh_ctx* c, int B, int64_t nblock, const int64_t* pos, int64_t npair, const int64_t* pair_hpos,
                     int64_t ndummy, const int64_t* dummy_pos) {
    if (!c || !c->prepared || (B != 3 && B != 6) || nblock < 1 || npair < 0 || ndummy < 0 || !pos ||
        (npair && !pair_hpos) || (ndummy && !dummy_pos))
        return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_chain(c);
    const int64_t npos = nblock * B;
    if (3 * npos * B >= INT32_MAX || 4 * npair >= INT32_MAX) {
        c->err = "chain system too large for 32-bit slots"; return MH_ERR_SIZE;
    }
    // every position is taken by exactly one dof or padding entry, every pair by one λ dof
    std::vector<int32_t> p32(c->nX), h32(npair), d32(ndummy);
    std::vector<char> used(npos, 0), lam(npair, 0);
    for (int64_t d = 0; d < c->nX; ++d) {
        const int64_t q = pos[d];
        if (q > 0 && q <= npos && !used[q - 1]) { used[q - 1] = 1; p32[d] = (int32_t)(q - 1); }
        else if (q < 0 && -q <= npair && !lam[-q - 1]) { lam[-q - 1] = 1; p32[d] = (int32_t)q; }
        else { c->err = "chain layout: bad or repeated position of X dof " + std::to_string(d + 1); return MH_ERR_ARG; }
    }
    for (int64_t q = 0; q < ndummy; ++q) {
        const int64_t v = dummy_pos[q];
        if (v < 1 || v > npos || used[v - 1]) { c->err = "chain layout: bad padding position"; return MH_ERR_ARG; }
        used[v - 1] = 1; d32[q] = (int32_t)(v - 1);
    }
    for (int64_t q = 0; q < npos; ++q)
        if (!used[q]) { c->err = "chain layout: position " + std::to_string(q + 1) + " is not taken"; return MH_ERR_ARG; }
    std::vector<char> held(npos, 0);
    for (int64_t p = 0; p < npair; ++p) {
        const int64_t v = pair_hpos[p];
        if (!lam[p] || v < 1 || v > npos || held[v - 1]) { c->err = "chain layout: bad λ pair " + std::to_string(p + 1); return MH_ERR_ARG; }
        held[v - 1] = 1; h32[p] = (int32_t)(v - 1);
    }
    const int64_t ntot = 3 * nblock * B * B + npos + 4 * npair;   // D | L | U | b | pair couplings
    HIPCHK(c, hipMalloc(&c->d_cpos, sizeof(int32_t) * std::max<int64_t>(c->nX, 1)));
    HIPCHK(c, hipMalloc(&c->d_chpos, sizeof(int32_t) * std::max<int64_t>(npair, 1)));
    HIPCHK(c, hipMalloc(&c->d_cdum, sizeof(int32_t) * std::max<int64_t>(ndummy, 1)));
    HIPCHK(c, hipMalloc(&c->d_cbuf, sizeof(double) * ntot));
    HIPCHK(c, hipMalloc(&c->d_csaved, sizeof(double) * std::max<int64_t>(npair * (3 * B + 1), 1)));
    HIPCHK(c, hipMalloc(&c->d_dx, sizeof(double) * std::max<int64_t>(c->nX, 1)));
    HIPCHK(c, hipMalloc(&c->d_cflag, chain::kStatusBytes));
    HIPCHK(c, hipHostMalloc((void**)&c->h_cflag, chain::kStatusBytes, hipHostMallocDefault));
    for (auto& e : c->cev) HIPCHK(c, hipEventCreate(&e));
    if (c->nX) HIPCHK(c, copy_sync(c, c->d_cpos, p32.data(), sizeof(int32_t) * c->nX, hipMemcpyHostToDevice));
    if (npair) HIPCHK(c, copy_sync(c, c->d_chpos, h32.data(), sizeof(int32_t) * npair, hipMemcpyHostToDevice));
    if (ndummy) HIPCHK(c, copy_sync(c, c->d_cdum, d32.data(), sizeof(int32_t) * ndummy, hipMemcpyHostToDevice));
    c->cB = B; c->cnb = nblock; c->cnp = npair; c->cnd = ndummy;
    ++c->gen;   // a captured Newton iteration holds the old chain buffers
    const int r = B == 6 ? build_chain_slots<6>(c) : build_chain_slots<3>(c);
    if (r != MH_OK) {
        const std::string e = c->err;
        free_chain(c);
        c->err = e;
    }
    return r;
}

// Reads the status words of the last chain solve (stream synchronized by the caller's copy).
static int chain_status_read(Ctx* c);
static int chain_status(Ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->h_cflag, c->d_cflag, chain::kStatusBytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return chain_status_read(c);
}
// The status words already in pinned host memory (after the stream has been synchronized).
static int chain_status_read(Ctx* c) {
    const unsigned f = c->h_cflag[0];
    const unsigned long long* where = reinterpret_cast<const unsigned long long*>(c->h_cflag + 2);
    if (f & chain::kBadPair) {
        c->err = "λ row with a nonzero diagonal or a zero coupling (pair " + std::to_string(where[1] + 1) + ")";
        return MH_ERR_SOLVE;
    }
    if (f & chain::kSingular) {
        const unsigned long long b = where[0];
        c->err = "zero or non-finite pivot in diagonal block " + std::to_string(b) + " (0-based, chain order; " +
                 (b ? "reduction level " + std::to_string(b & (~b + 1)) : std::string("the last block")) + ")";
        return MH_ERR_SOLVE;
    }
    return MH_OK;
}

int mh_chain_solve(mh_ctx* c) {
    if (!c || !c->prepared || !c->cB) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (int r = (c->cB == 6 ? enqueue_chain_solve<6>(c, c->d_rhs) : enqueue_chain_solve<3>(c, c->d_rhs))) return r;
    return chain_status(c);
}

// The adjoint solve once the unscaled ∂g/∂X is in d_crhs: scale, solve, Λ := alpha·scale ⊙ Δx.
static int adjoint_solve_tail(Ctx* c, double alpha) {
    hipStream_t s = c->stream;
    const int64_t n = std::max<int64_t>(c->nX, 1);
    hipLaunchKernelGGL(scaled_copy_kernel, dim3(grid(n)), dim3(kBlock), 0, s, c->nX, 1., c->d_scaleX, c->d_crhs, c->d_crhs);
    if (int r = (c->cB == 6 ? enqueue_chain_solve<6>(c, c->d_crhs) : enqueue_chain_solve<3>(c, c->d_crhs))) return r;
    hipLaunchKernelGGL(scaled_copy_kernel, dim3(grid(n)), dim3(kBlock), 0, s, c->nX, alpha, c->d_scaleX, c->d_dx, c->d_lam);
    if (int r = check_launch(c, "adjoint solve")) return r;
    return chain_status(c);
}
static int adjoint_buffers(Ctx* c) {
    const int64_t n = std::max<int64_t>(c->nX, 1);
    if (!c->d_crhs) HIPCHK(c, hipMalloc(&c->d_crhs, sizeof(double) * n));
    if (!c->d_lam) HIPCHK(c, hipMalloc(&c->d_lam, sizeof(double) * n));
    return MH_OK;
}

int mh_adjoint_solve(mh_ctx* c, const double* rhs, double alpha) {
    if (!c || !c->prepared || !c->cB || !rhs || !c->d_scaleX) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (int r = adjoint_buffers(c)) return r;
    if (c->nX) HIPCHK(c, hipMemcpyAsync(c->d_crhs, rhs, sizeof(double) * c->nX, hipMemcpyHostToDevice, c->stream));
    return adjoint_solve_tail(c, alpha);
}

int mh_adjoint_solve_sparse(mh_ctx* c, int64_t n, const int64_t* idx, const double* val, double alpha) {
    if (!c || !c->prepared || !c->cB || !c->d_scaleX || n < 0 || (n && (!idx || !val))) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    // nonzeros by dof, repeated dofs summed in their given order
    std::vector<std::pair<int64_t, int64_t>> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (idx[i] < 1 || idx[i] > c->nX) { c->err = "adjoint right-hand side: dof index out of range"; return MH_ERR_ARG; }
        order[(size_t)i] = {idx[i] - 1, i};
    }
    std::sort(order.begin(), order.end());
    std::vector<int64_t> hi;
    std::vector<double> hv;
    for (size_t k = 0; k < order.size(); ++k) {
        if (k && order[k].first == order[k - 1].first) hv.back() += val[order[k].second];
        else { hi.push_back(order[k].first); hv.push_back(val[order[k].second]); }
    }
    const int64_t m = (int64_t)hi.size();
    if (int r = adjoint_buffers(c)) return r;
    if (m > c->sp_cap) {
        dfree(c->d_spi); dfree(c->d_spv);
        c->sp_cap = std::max<int64_t>(m, 256);
        HIPCHK(c, hipMalloc(&c->d_spi, sizeof(int64_t) * c->sp_cap));
        HIPCHK(c, hipMalloc(&c->d_spv, sizeof(double) * c->sp_cap));
    }
    hipStream_t s = c->stream;
    const int64_t nx = std::max<int64_t>(c->nX, 1);
    hipLaunchKernelGGL(chain::fill_f64, dim3(cgrid(nx)), dim3(chain::kCB), 0, s, c->nX, c->d_crhs, 0.0);
    if (m) {
        HIPCHK(c, hipMemcpyAsync(c->d_spi, hi.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, s));
        HIPCHK(c, hipMemcpyAsync(c->d_spv, hv.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(sparse_set_kernel, dim3(grid(m)), dim3(kBlock), 0, s, m, c->d_spi, c->d_spv, c->d_crhs);
    }
    return adjoint_solve_tail(c, alpha);
}

int mh_chain_timings(mh_ctx* c, double* ms) {
    if (!c || !c->cB || !ms) return MH_ERR_ARG;
    float a = 0.f, b = 0.f, d = 0.f;
    HIPCHK(c, hipEventElapsedTime(&a, c->cev[0], c->cev[1]));
    HIPCHK(c, hipEventElapsedTime(&b, c->cev[1], c->cev[2]));
    HIPCHK(c, hipEventElapsedTime(&d, c->cev[2], c->cev[3]));
    ms[0] = a; ms[1] = b; ms[2] = d;
    return MH_OK;
}

int mh_get_dx(mh_ctx* c, double* dx) {
    if (!c || !c->cB || !dx) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nX) HIPCHK(c, copy_sync(c, dx, c->d_dx, sizeof(double) * c->nX, hipMemcpyDeviceToHost));
    return MH_OK;
}

int mh_newton_prepare(mh_ctx* c, const double* scaleX) {
    if (!c || (c->nX && !scaleX)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c->d_scaleX); dfree(c->d_part); dfree(c->d_norm);
    if (c->h_norm) { (void)hipHostFree(c->h_norm); c->h_norm = nullptr; }
    const int64_t n = std::max<int64_t>(c->nX, 1);
    HIPCHK(c, hipMalloc(&c->d_scaleX, sizeof(double) * n));
    HIPCHK(c, hipMalloc(&c->d_part, sizeof(double) * 2 * cgrid(n)));
    HIPCHK(c, hipMalloc(&c->d_norm, sizeof(double) * 2));
    HIPCHK(c, hipHostMalloc((void**)&c->h_norm, sizeof(double) * 2, hipHostMallocDefault));
    if (c->nX) HIPCHK(c, copy_sync(c, c->d_scaleX, scaleX, sizeof(double) * c->nX, hipMemcpyHostToDevice));
    ++c->gen;   // a captured Newton iteration holds the old buffers
    return MH_OK;
}

// Newmarkβdecrement!{OX} of the device X with the last Δx, and the norms into pinned memory
// (kernels and a copy on the stream: also nodes of a captured Newton iteration).
// The coefficients are read from d_c6 (upload_c6 first).
static int enqueue_newton_update(Ctx* c, int OX, int first) {
    hipStream_t s = c->stream;
    const int64_t ng = cgrid(std::max<int64_t>(c->nX, 1));
    hipLaunchKernelGGL(chain::newmark_kernel, dim3(ng), dim3(chain::kCB), 0, s, c->nX, OX, first, c->d_c6, c->d_dx,
                       c->d_rhs, c->d_scaleX, c->d_X[0], c->d_X[1], c->d_X[2], c->d_part);
    hipLaunchKernelGGL(chain::norm_kernel, dim3(1), dim3(chain::kCB), 0, s, ng, c->d_part, c->d_norm);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_norm, c->d_norm, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    return MH_OK;
}
static int check_newton_args(Ctx* c, int OX, const double* c6, const double* norms) {
    if (!c || !c->cB || !c->d_scaleX || OX < 0 || OX > 2 || (OX > 0 && !c6) || !norms) return MH_ERR_ARG;
    if (OX > 0 && !c->d_X[OX]) { c->err = "X derivatives of order " + std::to_string(OX) + " are not on the device"; return MH_ERR_ARG; }
    return MH_OK;
}

int mh_newton_update(mh_ctx* c, int OX, int first, const double* c6, double* norms) {
    if (int r = check_newton_args(c, OX, c6, norms)) return r;
    HIPCHK(c, hipSetDevice(c->device));
    if (int r = upload_c6(c, OX, c6)) return r;
    if (int r = enqueue_newton_update(c, OX, first)) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    norms[0] = c->h_norm[0];
    norms[1] = c->h_norm[1];
    return MH_OK;
}

int mh_newton_iteration_graph(mh_ctx* c, int OX, int first, const double* c6, double* norms, mh_err* err) {
    if (int r = check_sweepx_args(c, OX, (first && OX > 0) ? 0 : 1, c6, err)) return r;
    if (int r = check_newton_args(c, OX, c6, norms)) return r;
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->beam_tuned) {   // before any capture: the tuner times launches
        if (int r = tune_beams(c)) return r;
        ++c->gen;
    }
    const bool step = first && OX > 0;
    if (int r = upload_c6(c, OX, c6)) return r;
    const double key[9] = {0, 0, 0, 0, 0, 0, 0, 0, reads_time(c) ? c->t : 0.};
    const int sl = first ? 3 : 2;
    bool same = c->gexec[sl] && c->gOX[sl] == OX && c->ggen[sl] == c->gen;
    for (int i = 0; i < 9 && same; ++i) same = std::memcmp(&key[i], &c->gkey[sl][i], sizeof(double)) == 0;
    hipStream_t s = c->stream;
    if (!same) {
        free_graph(c, sl);
        HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int r = enqueue_sweepx(c, OX, step, c6, false, true);
        if (r == MH_OK)
            r = c->cB == 6 ? enqueue_chain_solve<6>(c, c->d_rhs, false, true) : enqueue_chain_solve<3>(c, c->d_rhs, false, true);
        if (r == MH_OK) {
            const hipError_t e = hipMemcpyAsync(c->h_cflag, c->d_cflag, chain::kStatusBytes, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) { c->err = std::string("capture: ") + hipGetErrorString(e); r = MH_ERR_HIP; }
        }
        if (r == MH_OK) r = enqueue_newton_update(c, OX, first);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (r != MH_OK || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (r == MH_OK) c->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec);
            return r != MH_OK ? r : MH_ERR_HIP;
        }
        c->graph[sl] = g;
        HIPCHK(c, hipGraphInstantiateWithFlags(&c->gexec[sl], c->graph[sl], 0));
        std::memcpy(c->gkey[sl], key, sizeof(key));
        c->gOX[sl] = OX; c->ggen[sl] = c->gen;
        ++c->graph_captures;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    HIPCHK(c, hipGraphLaunch(c->gexec[sl], s));
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    c->timed = true;          // mh_last_timings: the whole iteration
    c->graph_timed = true;
    if (int r = finish_sweepx(c, err)) return r;          // syncs the stream; NaN of the assembly
    if (int r = chain_status_read(c)) return r;           // pivots of the solve
    norms[0] = c->h_norm[0];
    norms[1] = c->h_norm[1];
    return MH_OK;
}

int mh_get_X(mh_ctx* c, int ider, double* x) {
    if (!c || ider < 0 || ider > 2 || !x || (c->nX && !c->d_X[ider])) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nX) HIPCHK(c, copy_sync(c, x, c->d_X[ider], sizeof(double) * c->nX, hipMemcpyDeviceToHost));
    return MH_OK;
}

}  // extern "C"

extern "C" {

int mh_beam_results(mh_ctx* c, int igroup, int ND, int ngauge, const double* gauges, double* out) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size() || ND < 1 || ND > 3 || ngauge < 0 || !out ||
        (ngauge > 0 && !gauges))
        return MH_ERR_ARG;
    const Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_BEAM && g.kind != MH_KIND_BEAM_U) { c->err = "group is not an EulerBeam3D group"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    if (g.nele == 0) return MH_OK;
    const int NR = ND == 1 ? mh::beam_nres(1) : ND == 2 ? mh::beam_nres(2) : mh::beam_nres(3);
    const int64_t n = g.nele * (NR + ngauge);
    double *d_out = nullptr, *d_g = nullptr;
    HIPCHK(c, hipMalloc(&d_out, sizeof(double) * n));
    if (ngauge) {
        if (hipMalloc(&d_g, sizeof(double) * 4 * ngauge * g.nele) != hipSuccess) {
            dfree(d_out); c->err = "hipMalloc"; return MH_ERR_HIP;
        }
    }
    BeamDynArgs b;
    b.nele = g.nele; b.idx = g.d_idx; b.uidx = g.d_uidx; b.nodes = g.d_nodes; b.xyz = c->d_xyz; b.o2 = g.d_o2;
    b.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) b.X[d] = c->d_X[d];
    b.U = c->d_U;
    hipStream_t s = c->stream;
    hipError_t e = hipSuccess;
    if (ngauge) e = hipMemcpyAsync(d_g, gauges, sizeof(double) * 4 * ngauge * g.nele, hipMemcpyHostToDevice, s);
    const dim3 gr((unsigned)((g.nele + kBlock - 1) / kBlock));
    if (e == hipSuccess) {
        if (ND == 1) hipLaunchKernelGGL(beam_results_kernel<1>, gr, dim3(kBlock), 0, s, b, ngauge, d_g, d_out);
        else if (ND == 2) hipLaunchKernelGGL(beam_results_kernel<2>, gr, dim3(kBlock), 0, s, b, ngauge, d_g, d_out);
        else hipLaunchKernelGGL(beam_results_kernel<3>, gr, dim3(kBlock), 0, s, b, ngauge, d_g, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_out); dfree(d_g);
    if (e != hipSuccess) { c->err = std::string("beam results: ") + hipGetErrorString(e); return MH_ERR_HIP; }
    return MH_OK;
}

}  // extern "C"

extern "C" {

int mh_beam_gauge_derivs(mh_ctx* c, int igroup, int64_t nsel, const int64_t* sel, double* out) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size() || nsel < 0 || !out) return MH_ERR_ARG;
    const Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_BEAM && g.kind != MH_KIND_BEAM_U) { c->err = "group is not an EulerBeam3D group"; return MH_ERR_ARG; }
    const int64_t n = sel ? nsel : g.nele;
    if (sel)
        for (int64_t k = 0; k < nsel; ++k)
            if (sel[k] < 1 || sel[k] > g.nele) { c->err = "element number out of range"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    if (n == 0) return MH_OK;
    std::vector<int32_t> s32;
    if (sel) { s32.resize(n); for (int64_t k = 0; k < n; ++k) s32[k] = (int32_t)(sel[k] - 1); }
    double* d_out = nullptr;
    int32_t* d_sel = nullptr;
    HIPCHK(c, hipMalloc(&d_out, sizeof(double) * n * kGaugeOut));
    if (sel && hipMalloc(&d_sel, sizeof(int32_t) * n) != hipSuccess) { dfree(d_out); c->err = "hipMalloc"; return MH_ERR_HIP; }
    BeamDynArgs b;
    b.nele = g.nele; b.idx = g.d_idx; b.uidx = g.d_uidx; b.nodes = g.d_nodes; b.xyz = c->d_xyz; b.o2 = g.d_o2;
    b.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) b.X[d] = c->d_X[d];
    b.U = c->d_U;
    hipStream_t s = c->stream;
    hipError_t e = hipSuccess;
    if (sel) e = hipMemcpyAsync(d_sel, s32.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL((beam_gauge_derivs_kernel<int32_t, 0>), dim3((unsigned)((n * kPairs + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           b, d_sel, n, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * n * kGaugeOut, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_out); dfree(d_sel);
    if (e != hipSuccess) { c->err = std::string("gauge derivatives: ") + hipGetErrorString(e); return MH_ERR_HIP; }
    return MH_OK;
}

int mh_beam_dRdA(mh_ctx* c, int igroup, double* out) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size() || !out) return MH_ERR_ARG;
    const Group& g = c->groups[igroup];
    if (!g.amask) { c->err = "group has no A-scaled stiffness (mh_group_set_adofs)"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    if (g.nele == 0) return MH_OK;
    double* d_out = nullptr;
    HIPCHK(c, hipMalloc(&d_out, sizeof(double) * 12 * g.nele));
    VjpArgs a{};
    a.nele = g.nele; a.nX = g.nX; a.kind = g.kind; a.idx = g.d_idx; a.uidx = g.d_uidx; a.nodes = g.d_nodes;
    a.xyz = c->d_xyz; a.o2 = g.d_o2; a.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) a.X[d] = c->d_X[d];
    a.U = c->d_U; a.lam = nullptr; a.flag = c->d_flag; a.gtag = (unsigned long long)igroup << 40;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, s, c->d_flag, ~0ull);
    hipLaunchKernelGGL(beam_vjp_a_kernel<true>, dim3((unsigned)((g.nele + kBeamBlock - 1) / kBeamBlock)), dim3(kBeamBlock),
                       0, s, a, g.amask, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * 12 * g.nele, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(d_out);
    if (e != hipSuccess) { c->err = std::string("dR/dA: ") + hipGetErrorString(e); return MH_ERR_HIP; }
    return finish_sweepx(c, nullptr);
}

// mh_beam_gauge_derivs without any allocation: the selection and the output staging live in
// this device's fixed scratch array (kGaugeNoAllocMax elements at most); the kernel is the same.
int mh_beam_gauge_derivs_noalloc(mh_ctx* c, int igroup, int64_t nsel, const int64_t* sel, double* out) {
    if (!c || igroup < 0 || igroup >= (int)c->groups.size() || nsel < 0 || !out) return MH_ERR_ARG;
    const Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_BEAM && g.kind != MH_KIND_BEAM_U) { c->err = "group is not an EulerBeam3D group"; return MH_ERR_ARG; }
    const int64_t n = sel ? nsel : g.nele;
    if (n > kGaugeNoAllocMax) {
        c->err = "gauge derivatives without allocation: " + std::to_string(n) + " elements requested, at most " +
                 std::to_string(kGaugeNoAllocMax) + " per call (the fixed scratch)";
        return MH_ERR_SIZE;
    }
    if (!c->d_scratch) { c->err = "no fixed scratch on this device"; return MH_ERR_ARG; }
    if (sel)
        for (int64_t k = 0; k < nsel; ++k)
            if (sel[k] < 1 || sel[k] > g.nele) { c->err = "element number out of range"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    if (n == 0) return MH_OK;
    int64_t* d_sel = static_cast<int64_t*>(c->d_scratch);
    double* d_out = reinterpret_cast<double*>(static_cast<unsigned char*>(c->d_scratch) + kScratchSelBytes);
    BeamDynArgs b;
    b.nele = g.nele; b.idx = g.d_idx; b.uidx = g.d_uidx; b.nodes = g.d_nodes; b.xyz = c->d_xyz; b.o2 = g.d_o2;
    b.sec = g.d_sec;
    for (int d = 0; d < 3; ++d) b.X[d] = c->d_X[d];
    b.U = c->d_U;
    hipStream_t s = c->stream;
    hipError_t e = hipSuccess;
    if (sel) e = hipMemcpyAsync(d_sel, sel, sizeof(int64_t) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL((beam_gauge_derivs_kernel<int64_t, 1>), dim3((unsigned)((n * kPairs + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, b, sel ? d_sel : nullptr, n, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * n * kGaugeOut, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { c->err = std::string("gauge derivatives: ") + hipGetErrorString(e); return MH_ERR_HIP; }
    return MH_OK;
}

}  // extern "C"

extern "C" {

int mh_set_host_blocks(mh_ctx* c, int igroup, const double* R, const double* K) {
    if (!c || !c->prepared || igroup < 0 || igroup >= (int)c->groups.size() || !R || !K) return MH_ERR_ARG;
    const Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_HOST) { c->err = "group is not host-evaluated"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = g.nele;
    const int nX = g.nX;
    if (n == 0) return MH_OK;
    std::vector<double> r((size_t)nX * n), k((size_t)nX * nX * n);
    for (int64_t e = 0; e < n; ++e)
        for (int i = 0; i < nX; ++i) {
            r[(size_t)i * n + e] = R[(size_t)nX * e + i] * g.scaleX[i];
            for (int j = 0; j < nX; ++j) {
                const size_t q = (size_t)nX * nX * e + i + (size_t)nX * j;
                k[q] = K[q] * g.scaleX[i] * g.scaleX[j];
            }
        }
    HIPCHK(c, hipMemcpyAsync(c->d_Rbuf + g.rbase, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_Kbuf + g.kbase, k.data(), sizeof(double) * k.size(), hipMemcpyHostToDevice, c->stream));
    // a static element has no ∂R/∂δr: the :step rhs term is zero until mh_set_host_dr gives one
    std::fill(r.begin(), r.end(), 0.);
    HIPCHK(c, hipMemcpyAsync(c->d_dRbuf + g.rbase, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}

int mh_set_host_dr(mh_ctx* c, int igroup, const double* dR) {
    if (!c || !c->prepared || igroup < 0 || igroup >= (int)c->groups.size() || !dR) return MH_ERR_ARG;
    const Group& g = c->groups[igroup];
    if (g.kind != MH_KIND_HOST) { c->err = "group is not host-evaluated"; return MH_ERR_ARG; }
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = g.nele;
    const int nX = g.nX;
    if (n == 0) return MH_OK;
    std::vector<double> r((size_t)nX * n);
    for (int64_t e = 0; e < n; ++e)
        for (int i = 0; i < nX; ++i) r[(size_t)i * n + e] = dR[(size_t)nX * e + i] * g.scaleX[i];
    HIPCHK(c, hipMemcpyAsync(c->d_dRbuf + g.rbase, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MH_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ DirectXUA big system
namespace {

__global__ __launch_bounds__(kBlock) void iota_u32(int64_t n, uint32_t* v) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) v[p] = (uint32_t)p;
}
__global__ __launch_bounds__(kBlock) void big_permute(int64_t n, const uint32_t* perm, const int64_t* src_in,
                                                      const double* w_in, int64_t* src, double* w) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t q = perm[p];
    src[p] = src_in[q];
    w[p] = w_in[q];
}
// val[k] = Σ w·source over the contributions of nz k, in contribution order (deterministic).
__global__ __launch_bounds__(kBlock) void big_scatter(int64_t nnz, const int32_t* __restrict__ seg,
                                                      const int64_t* __restrict__ src, const double* __restrict__ w,
                                                      const double* __restrict__ dev, int64_t ndev,
                                                      const double* __restrict__ host, double* __restrict__ val) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    double acc = 0.;
    for (int32_t p = seg[k]; p < seg[k + 1]; ++p) {
        const int64_t q = src[p];
        acc += w[p] * (q < ndev ? dev[q] : host[q - ndev]);
    }
    val[k] = acc;
}

}  // namespace

extern "C" {

int mh_big_prepare(mh_ctx* c, int64_t nv, int64_t np, const int64_t* rows, const int64_t* cols, const int64_t* src,
                   const double* w, int64_t nhost, int64_t* nnz_out) {
    if (!c || !c->direct || nv < 1 || np < 0 || nhost < 0 || (np && (!rows || !cols || !src || !w))) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    free_big(c);
    if (nv >= INT32_MAX || np >= INT32_MAX) { c->err = "big system too large for 32-bit maps"; return MH_ERR_SIZE; }
    const int64_t ndev = 2 * (int64_t)(c->dOX + 1) * c->dXX.nnz + 2 * c->dXU.nnz;
    std::vector<unsigned long long> keys(np);
    for (int64_t p = 0; p < np; ++p) {
        if (rows[p] < 0 || rows[p] >= nv || cols[p] < 0 || cols[p] >= nv || src[p] < 0 || src[p] >= ndev + nhost) {
            c->err = "big contribution " + std::to_string(p + 1) + " out of range"; return MH_ERR_ARG;
        }
        keys[p] = (unsigned long long)cols[p] * (unsigned long long)nv + (unsigned long long)rows[p];
    }
    hipStream_t s = c->stream;
    c->bnv = nv; c->bnp = np; c->bnhost = nhost; c->bndev = ndev;
    HIPCHK(c, hipMalloc(&c->d_bcolptr, sizeof(int32_t) * (nv + 1)));
    HIPCHK(c, hipMalloc(&c->d_bhost, sizeof(double) * std::max<int64_t>(nhost, 1)));
    if (np == 0) {
        HIPCHK(c, hipMemsetAsync(c->d_bcolptr, 0, sizeof(int32_t) * (nv + 1), s));
        HIPCHK(c, hipStreamSynchronize(s));
        if (nnz_out) *nnz_out = 0;
        return MH_OK;
    }
    unsigned long long *kin = nullptr, *kout = nullptr;
    uint32_t *vin = nullptr, *perm = nullptr;
    int32_t *head = nullptr, *id = nullptr, *colk = nullptr;
    int64_t* src_in = nullptr;
    double* w_in = nullptr;
    HIPCHK(c, hipMalloc(&kin, sizeof(unsigned long long) * np));
    HIPCHK(c, hipMalloc(&kout, sizeof(unsigned long long) * np));
    HIPCHK(c, hipMalloc(&vin, sizeof(uint32_t) * np));
    HIPCHK(c, hipMalloc(&perm, sizeof(uint32_t) * np));
    HIPCHK(c, hipMalloc(&src_in, sizeof(int64_t) * np));
    HIPCHK(c, hipMalloc(&w_in, sizeof(double) * np));
    HIPCHK(c, hipMemcpyAsync(kin, keys.data(), sizeof(unsigned long long) * np, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(src_in, src, sizeof(int64_t) * np, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(w_in, w, sizeof(double) * np, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(iota_u32, dim3(grid(np)), dim3(kBlock), 0, s, np, vin);
    int end_bit = 1;
    while (end_bit < 64 && ((unsigned long long)nv * (unsigned long long)nv) > (1ull << end_bit)) ++end_bit;
    size_t tmp_bytes = 0;
    void* tmp = nullptr;
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kin, kout, vin, perm, (int)np, 0, end_bit, s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, perm, (int)np, 0, end_bit, s));
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(tmp);
    HIPCHK(c, hipMalloc(&head, sizeof(int32_t) * np));
    HIPCHK(c, hipMalloc(&id, sizeof(int32_t) * np));
    hipLaunchKernelGGL(head_kernel<unsigned long long>, dim3(grid(np)), dim3(kBlock), 0, s, np, kout, head);
    tmp_bytes = 0;
    HIPCHK(c, hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, head, id, (int)np, s));
    HIPCHK(c, hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
    HIPCHK(c, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, head, id, (int)np, s));
    int32_t nnz = 0;
    HIPCHK(c, hipMemcpyAsync(&nnz, id + np - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(tmp);
    c->bnnz = nnz;
    HIPCHK(c, hipMalloc(&c->d_browval, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    HIPCHK(c, hipMalloc(&c->d_bseg, sizeof(int32_t) * (nnz + 1)));
    HIPCHK(c, hipMalloc(&c->d_bsrc, sizeof(int64_t) * np));
    HIPCHK(c, hipMalloc(&c->d_bw, sizeof(double) * np));
    HIPCHK(c, hipMalloc(&c->d_bval, sizeof(double) * std::max<int64_t>(nnz, 1)));
    HIPCHK(c, hipMalloc(&colk, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    hipLaunchKernelGGL(seg_kernel, dim3(grid(np)), dim3(kBlock), 0, s, np, kout, head, id, nv, c->d_bseg, c->d_browval, colk);
    const int32_t np32 = (int32_t)np;
    HIPCHK(c, hipMemcpyAsync(c->d_bseg + nnz, &np32, sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gapfill_kernel, dim3(grid(nnz + 1)), dim3(kBlock), 0, s, (int64_t)nnz, colk, nv, c->d_bcolptr);
    hipLaunchKernelGGL(big_permute, dim3(grid(np)), dim3(kBlock), 0, s, np, perm, src_in, w_in, c->d_bsrc, c->d_bw);
    if (int r = check_launch(c, "big pattern")) return r;
    HIPCHK(c, hipStreamSynchronize(s));
    dfree(kin); dfree(kout); dfree(vin); dfree(perm); dfree(head); dfree(id); dfree(colk); dfree(src_in); dfree(w_in);
    if (nnz_out) *nnz_out = nnz;
    return MH_OK;
}

int mh_big_get_pattern(mh_ctx* c, int64_t* colptr, int64_t* rowval) {
    if (!c || !c->d_bcolptr || !colptr || (c->bnnz && !rowval)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int32_t> cp(c->bnv + 1), rv(c->bnnz);
    HIPCHK(c, copy_sync(c, cp.data(), c->d_bcolptr, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost));
    if (c->bnnz) HIPCHK(c, copy_sync(c, rv.data(), c->d_browval, sizeof(int32_t) * rv.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < cp.size(); ++i) colptr[i] = (int64_t)cp[i] + 1;
    for (size_t i = 0; i < rv.size(); ++i) rowval[i] = (int64_t)rv[i] + 1;
    return MH_OK;
}

int mh_big_assemble(mh_ctx* c, const double* hostvals, double* nzval) {
    if (!c || !c->direct || !c->d_bcolptr || (c->bnhost && !hostvals)) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (c->bnhost) HIPCHK(c, hipMemcpyAsync(c->d_bhost, hostvals, sizeof(double) * c->bnhost, hipMemcpyHostToDevice, s));
    if (c->bnnz) {
        hipLaunchKernelGGL(big_scatter, dim3(grid(c->bnnz)), dim3(kBlock), 0, s, c->bnnz, c->d_bseg, c->d_bsrc, c->d_bw,
                           c->d_dL2, c->bndev, c->d_bhost, c->d_bval);
        if (int r = check_launch(c, "big scatter")) return r;
        if (nzval) HIPCHK(c, hipMemcpyAsync(nzval, c->d_bval, sizeof(double) * c->bnnz, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(c, hipStreamSynchronize(s));
    return MH_OK;
}

#define BLASCHK(ctx, call)                                                                      \
    do {                                                                                        \
        rocblas_status st_ = (call);                                                            \
        if (st_ != rocblas_status_success) {                                                    \
            (ctx)->err = std::string(#call) + ": " + rocblas_status_to_string(st_);              \
            return MH_ERR_HIP;                                                                  \
        }                                                                                       \
    } while (0)

// Dense block storage of the all-steps system, LAPACK band storage at block level: block column j
// (m columns) holds block rows j−2 (fill from the row interchanges), j−1, j, j+1 as one 4m × m
// column-major panel, so that the band is one 4m × (nb·m) matrix with leading dimension 4m and
// block (i, j) starts at row (i − j + 2)·m. W (nb·m × (na+1), leading dimension nb·m) holds the
// border columns of the step rows and the right-hand side (later the solution); C (na × nb·m) the
// border rows; E (na × (na+1)) the corner and the border's right-hand side.
__device__ __forceinline__ int64_t bsolve_vpos(int64_t i, int64_t nsp, int64_t nbm, int64_t na, int64_t oW,
                                               int64_t oE) {
    return i < nsp ? oW + nbm * na + i : oE + na * na + (i - nsp);
}
// LU with partial pivoting of a leaf panel (M ≤ kPT·kPR rows × w ≤ kPW columns, leading
// dimension lda) in one workgroup, the steps of LAPACK's dgetf2: the first largest |entry| as
// pivot, the row interchange, the column scaled by the pivot's reciprocal, the update. The panel
// goes in sub-panels of kSW columns held in registers (row i in thread i % kPT, slot i / kPT);
// after each sub-panel its rows of the columns to the right are solved with its unit lower
// triangle and the rows below take the rank-kSW update, in global memory. piv[j] = roff + pivot
// row + 1; *info = roff + j + 1 at the first exactly zero pivot (the column is then left as is).
constexpr int kPW = 32;
constexpr int kPT = 1024;
constexpr int kPR = 5;
constexpr int kSW = 4;
__global__ __launch_bounds__(kPT) void panel_lu_kernel(int M, int w, double* __restrict__ P, int lda,
                                                       int32_t* __restrict__ piv, int32_t* __restrict__ info, int roff) {
    __shared__ double sv[kPT / 64];
    __shared__ int si[kPT / 64];
    __shared__ int sp;
    __shared__ double prow[kSW], jrow[kSW];
    __shared__ double L11[kSW][kSW];
    __shared__ double U12[kSW][kPW];
    const int t = threadIdx.x;
    for (int sb = 0; sb < w; sb += kSW) {
        const int sw = min(kSW, w - sb);
        double v[kPR][kSW];
#pragma unroll
        for (int r = 0; r < kPR; ++r) {
            const int i = t + r * kPT;
#pragma unroll
            for (int c = 0; c < kSW; ++c) v[r][c] = (i < M && c < sw) ? P[(int64_t)(sb + c) * lda + i] : 0.;
        }
#pragma unroll
        for (int jj = 0; jj < kSW; ++jj) {
            if (jj >= sw) break;
            const int j = sb + jj;
            double best = -1.;
            int bi = M;
#pragma unroll
            for (int r = 0; r < kPR; ++r) {   // a thread's rows increase with r: > keeps the first
                const int i = t + r * kPT;
                const double a = fabs(v[r][jj]);
                if (i >= j && i < M && a > best) { best = a; bi = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if ((t & 63) == 0) { sv[t >> 6] = best; si[t >> 6] = bi; }
            __syncthreads();
            if (t == 0) {
                double b = sv[0];
                int i0 = si[0];
                for (int q = 1; q < kPT / 64; ++q)
                    if (sv[q] > b || (sv[q] == b && si[q] < i0)) { b = sv[q]; i0 = si[q]; }
                if (i0 >= M) i0 = j;   // nothing comparable in the column (NaN): keep the row
                sp = i0;
                piv[j] = roff + i0 + 1;
            }
            __syncthreads();
            const int p = sp;
            const int tp = p % kPT, rp = p / kPT;
            if (t == tp) {
#pragma unroll
                for (int r = 0; r < kPR; ++r)
                    if (r == rp) {
#pragma unroll
                        for (int c = 0; c < kSW; ++c) prow[c] = v[r][c];
                    }
            }
            if (t == j) {   // row j < kPW lives in thread j, slot 0
#pragma unroll
                for (int c = 0; c < kSW; ++c) jrow[c] = v[0][c];
            }
            // the same interchange in the panel's other columns (one thread per column, in order)
            if (p != j && t < w && (t < sb || t >= sb + sw)) {
                double* cc = P + (int64_t)t * lda;
                const double x = cc[j];
                cc[j] = cc[p];
                cc[p] = x;
            }
            __syncthreads();
            if (p != j) {
                if (t == j) {
#pragma unroll
                    for (int c = 0; c < kSW; ++c) v[0][c] = prow[c];
                }
                if (t == tp) {
#pragma unroll
                    for (int r = 0; r < kPR; ++r)
                        if (r == rp) {
#pragma unroll
                            for (int c = 0; c < kSW; ++c) v[r][c] = jrow[c];
                        }
                }
            }
            const double d = prow[jj];
            if (d == 0.) {
                if (t == 0 && *info == 0) *info = roff + j + 1;
            } else {
                const double inv = 1. / d;
#pragma unroll
                for (int r = 0; r < kPR; ++r) {
                    const int i = t + r * kPT;
                    if (i > j && i < M) {
                        const double l = v[r][jj] * inv;
                        v[r][jj] = l;
#pragma unroll
                        for (int c = jj + 1; c < kSW; ++c) v[r][c] -= l * prow[c];
                    }
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < kPR; ++r) {
            const int i = t + r * kPT;
            if (i < M) {
#pragma unroll
                for (int c = 0; c < kSW; ++c)
                    if (c < sw) P[(int64_t)(sb + c) * lda + i] = v[r][c];
            }
        }
        if (t >= sb && t < sb + sw) {   // rows sb.. of the sub-panel: thread = row, slot 0
#pragma unroll
            for (int c = 0; c < kSW; ++c) L11[t - sb][c] = v[0][c];
        }
        __syncthreads();
        const int nr = w - sb - sw;
        if (nr > 0) {
            if (t < nr) {   // U12 = L11⁻¹ A12 for right column sb + sw + t
                double* cc = P + (int64_t)(sb + sw + t) * lda + sb;
                double u[kSW];
#pragma unroll
                for (int q = 0; q < kSW; ++q) u[q] = q < sw ? cc[q] : 0.;
#pragma unroll
                for (int q = 1; q < kSW; ++q)
#pragma unroll
                    for (int k = 0; k < q; ++k) u[q] -= L11[q][k] * u[k];
#pragma unroll
                for (int q = 0; q < kSW; ++q) {
                    if (q < sw) cc[q] = u[q];
                    U12[q][t] = q < sw ? u[q] : 0.;
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kPR; ++r) {
                const int i = t + r * kPT;
                if (i >= sb + sw && i < M) {
                    double* row = P + (int64_t)(sb + sw) * lda + i;
#pragma unroll 4
                    for (int c = 0; c < nr; ++c) {
                        double a = row[(int64_t)c * lda];
#pragma unroll
                        for (int q = 0; q < kSW; ++q) a -= v[r][q] * U12[q][c];
                        row[(int64_t)c * lda] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}
// Row interchanges piv[0..n) (1-based, offset roff) applied in order to ncol columns of A, one
// thread per column (for few interchanges).
__global__ __launch_bounds__(kBlock) void laswp_cols(int ncol, double* __restrict__ A, int lda, int n,
                                                     const int32_t* __restrict__ piv, int roff) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= ncol) return;
    double* a = A + (int64_t)q * lda;
    for (int i = 0; i < n; ++i) {
        const int p = piv[i] - 1 - roff;
        if (p != i) { const double x = a[i]; a[i] = a[p]; a[p] = x; }
    }
}
// The row permutation of the interchanges piv[0..n) (1-based, offset roff) over `rows` rows: row i
// of the result is row perm[i] of the original. One workgroup, in LDS.
__global__ __launch_bounds__(256) void piv_to_perm(int n, int rows, const int32_t* __restrict__ piv, int roff,
                                                   int32_t* __restrict__ perm) {
    extern __shared__ int32_t shp[];
    int32_t* pr = shp;
    int32_t* pv = shp + rows;
    for (int i = threadIdx.x; i < rows; i += 256) pr[i] = i;
    for (int i = threadIdx.x; i < n; i += 256) pv[i] = piv[i] - 1 - roff;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < n; ++i) {
            const int p = pv[i];
            if (p != i) { const int32_t x = pr[i]; pr[i] = pr[p]; pr[p] = x; }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < rows; i += 256) perm[i] = pr[i];
}
// Rows of each column permuted: a[i] ← a[perm[i]], one workgroup per column through LDS.
__global__ __launch_bounds__(256) void perm_rows(int rows, double* __restrict__ A, int lda, const int32_t* __restrict__ perm) {
    extern __shared__ double shc[];
    double* a = A + (int64_t)blockIdx.x * lda;
    for (int i = threadIdx.x; i < rows; i += 256) shc[i] = a[i];
    __syncthreads();
    for (int i = threadIdx.x; i < rows; i += 256) a[i] = shc[perm[i]];
}

// Equilibration (as SuperLU's default): row scales sr = 1/max|row|, then column scales
// sc = 1/max|sr·column|; the dense factors hold diag(sr)·M·diag(sc).
__global__ __launch_bounds__(kBlock) void bsolve_rowscale(int64_t nv, const int32_t* __restrict__ rowptr,
                                                          const int32_t* __restrict__ perm, const double* __restrict__ val,
                                                          double* __restrict__ sr) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    double mx = 0.;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) mx = fmax(mx, fabs(val[perm[p]]));
    sr[i] = mx > 0. ? 1. / mx : 1.;
}
__global__ __launch_bounds__(kBlock) void bsolve_colscale(int64_t nv, const int32_t* __restrict__ colptr,
                                                          const int32_t* __restrict__ rowval, const double* __restrict__ val,
                                                          const double* __restrict__ sr, double* __restrict__ sc) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= nv) return;
    double mx = 0.;
    for (int32_t p = colptr[j]; p < colptr[j + 1]; ++p) mx = fmax(mx, fabs(val[p] * sr[rowval[p]]));
    sc[j] = mx > 0. ? 1. / mx : 1.;
}
__global__ __launch_bounds__(kBlock) void bsolve_fill(int64_t nnz, const int64_t* __restrict__ slot,
                                                      const double* __restrict__ val, const int32_t* __restrict__ rowval,
                                                      const int32_t* __restrict__ colv, const double* __restrict__ sr,
                                                      const double* __restrict__ sc, double* __restrict__ buf) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < nnz) buf[slot[k]] = (val[k] * sr[rowval[k]]) * sc[colv[k]];
}
__global__ __launch_bounds__(kBlock) void bsolve_ones(int64_t n, const int64_t* __restrict__ pos, double* __restrict__ buf) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < n) buf[pos[k]] = 1.;
}
__global__ __launch_bounds__(kBlock) void bsolve_rhs(int64_t nv, int64_t nsp, int64_t nbm, int64_t na, int64_t oW,
                                                     int64_t oE, const double* __restrict__ rhs,
                                                     const double* __restrict__ sr, double* __restrict__ buf) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nv) buf[bsolve_vpos(i, nsp, nbm, na, oW, oE)] = rhs[i] * sr[i];
}
// r = b − M x over the pattern by rows, each row summed in column order (deterministic), and
// each row's componentwise backward error |r_i| / (Σ_j |M_ij x_j| + |b_i|) (∞ for a nonzero
// residual over a zero denominator)
__global__ __launch_bounds__(kBlock) void bsolve_resid(int64_t nv, const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ colidx, const int32_t* __restrict__ perm,
                                                       const double* __restrict__ val, const double* __restrict__ x,
                                                       const double* __restrict__ b, double* __restrict__ r,
                                                       double* __restrict__ wrow) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    double acc = 0., den = fabs(b[i]);
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const double t = val[perm[p]] * x[colidx[p]];
        acc += t;
        den += fabs(t);
    }
    const double ri = b[i] - acc;
    r[i] = ri;
    const double q = fabs(ri);
    wrow[i] = den > 0. ? q / den : (q > 0. ? INFINITY : 0.);
}
// ω = max_i wrow_i, the componentwise backward error of x. One workgroup.
__global__ __launch_bounds__(kBlock) void bsolve_omega(int64_t nv, const double* __restrict__ wrow,
                                                       double* __restrict__ omega) {
    __shared__ double red[kBlock];
    double w = 0.;
    for (int64_t i = threadIdx.x; i < nv; i += kBlock) w = fmax(w, wrow[i]);
    red[threadIdx.x] = w;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *omega = red[0];
}
__global__ __launch_bounds__(kBlock) void bsolve_add(int64_t nv, const double* __restrict__ d, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nv) x[i] += d[i];
}
__global__ __launch_bounds__(kBlock) void bsolve_sol(int64_t nv, int64_t nsp, int64_t nbm, int64_t na, int64_t oW,
                                                     int64_t oE, const double* __restrict__ buf,
                                                     const double* __restrict__ sc, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < nv) x[i] = buf[bsolve_vpos(i, nsp, nbm, na, oW, oE)] * sc[i];
}

// Recursive LU with partial pivoting (LAPACK's getrf2 splitting) of an M × N block with leading
// dimension lda: halves down to leaf panels of kPW columns (panel_lu_kernel), the row
// interchanges composed into permutations where they are many, rocBLAS trsm and gemm for the
// off-diagonal parts. piv holds global 1-based rows (offset roff = the block's first row).
struct RecLU {
    Ctx* c;
    hipStream_t s;
    rocblas_handle h;
    int lda;
    int32_t* info;
    int32_t* tmp;   // a row permutation (M entries)
    int swap(double* A, int M, int ncol, int n, const int32_t* piv, int roff) {
        if (ncol <= 0 || n <= 0) return MH_OK;
        if (n <= 64) {
            hipLaunchKernelGGL(laswp_cols, dim3((ncol + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ncol, A, lda, n, piv, roff);
        } else {
            hipLaunchKernelGGL(piv_to_perm, dim3(1), dim3(256), sizeof(int32_t) * (M + n), s, n, M, piv, roff, tmp);
            hipLaunchKernelGGL(perm_rows, dim3(ncol), dim3(256), sizeof(double) * M, s, M, A, lda, tmp);
        }
        return MH_OK;
    }
    int run(double* A, int M, int N, int32_t* piv, int roff) {
        if (N <= kPW) {
            hipLaunchKernelGGL(panel_lu_kernel, dim3(1), dim3(kPT), 0, s, M, N, A, lda, piv, info, roff);
            return MH_OK;
        }
        const int n1 = ((N / 2) + kPW - 1) / kPW * kPW;
        const double one = 1., mone = -1.;
        const rocblas_operation Nop = rocblas_operation_none;
        if (int r = run(A, M, n1, piv, roff)) return r;
        swap(A + (int64_t)n1 * lda, M, N - n1, n1, piv, roff);
        BLASCHK(c, rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, Nop, rocblas_diagonal_unit, n1, N - n1, &one, A,
                                 lda, A + (int64_t)n1 * lda, lda));
        BLASCHK(c, rocblas_dgemm(h, Nop, Nop, M - n1, N - n1, n1, &mone, A + n1, lda, A + (int64_t)n1 * lda, lda, &one,
                                 A + n1 + (int64_t)n1 * lda, lda));
        if (int r = run(A + n1 + (int64_t)n1 * lda, M - n1, N - n1, piv + n1, roff + n1)) return r;
        swap(A + n1, M - n1, n1, N - n1, piv + n1, roff + n1);
        return MH_OK;
    }
};

int mh_big_solve_prepare(mh_ctx* c, int64_t per_step, int64_t nstep, int64_t na, int64_t* m_out, int64_t* nb_out) {
    if (!c || !c->d_bcolptr || per_step < 1 || nstep < 1 || na < 0 || per_step * nstep + na != c->bnv) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    HIPCHK(c, hipStreamSynchronize(s));
    free_bsolve(c);
    const int64_t nv = c->bnv, nnz = c->bnnz, nsp = per_step * nstep;
    std::vector<int32_t> cp(nv + 1), rv(std::max<int64_t>(nnz, 1));
    HIPCHK(c, copy_sync(c, cp.data(), c->d_bcolptr, sizeof(int32_t) * (nv + 1), hipMemcpyDeviceToHost));
    if (nnz) HIPCHK(c, copy_sync(c, rv.data(), c->d_browval, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    // the widest coupling between steps fixes how many steps one diagonal block holds, so that the
    // blocks are tridiagonal (finite differences of order OX couple steps up to two apart)
    int64_t bw = 0;
    for (int64_t j = 0; j < nsp; ++j)
        for (int32_t p = cp[j]; p < cp[j + 1]; ++p) {
            const int64_t r = rv[p];
            if (r < nsp) bw = std::max<int64_t>(bw, std::llabs(r / per_step - j / per_step));
        }
    const int64_t g = std::max<int64_t>(bw, 1), m = g * per_step, nb = (nstep + g - 1) / g, nbm = nb * m;
    // the row permutations go through LDS: 2m doubles per column, 3m indices
    if (4 * m >= INT32_MAX / 2 || nbm >= INT32_MAX / 2 || na >= 65536 || 2 * m * 8 > 64000) {
        c->err = "diagonal blocks too large for a dense solve"; return MH_ERR_SIZE;
    }
    int64_t* o = c->so;
    o[0] = 0; o[1] = 4 * m * nbm; o[2] = o[1] + nbm * (na + 1); o[3] = o[2] + na * nbm;
    const int64_t tot = o[3] + na * (na + 1);
    size_t fr = 0, tt = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tt));
    if ((double)tot * 8. > 0.7 * (double)fr) {
        c->err = "the dense block storage (" + std::to_string(tot * 8 / (1 << 20)) + " MiB) exceeds the free device memory";
        return MH_ERR_SIZE;
    }
    std::vector<int64_t> slot(std::max<int64_t>(nnz, 1));
    for (int64_t j = 0; j < nv; ++j)
        for (int32_t p = cp[j]; p < cp[j + 1]; ++p) {
            const int64_t r = rv[p];
            int64_t q;
            if (r < nsp && j < nsp) {
                const int64_t kr = (r / per_step) / g, kc = (j / per_step) / g;
                q = o[0] + j * 4 * m + (kr - kc + 2) * m + (r - kr * m);
            } else if (r < nsp) {
                q = o[1] + r + nbm * (j - nsp);
            } else if (j < nsp) {
                q = o[2] + (r - nsp) + na * j;
            } else {
                q = o[3] + (r - nsp) + na * (j - nsp);
            }
            slot[p] = q;
        }
    std::vector<int64_t> pad;   // identity rows for the unused rows of the last block
    for (int64_t i = nsp; i < nbm; ++i) pad.push_back(o[0] + i * 4 * m + 2 * m + (i - (nb - 1) * m));
    HIPCHK(c, hipMalloc(&c->d_sslot, sizeof(int64_t) * slot.size()));
    HIPCHK(c, hipMalloc(&c->d_spad, sizeof(int64_t) * std::max<size_t>(pad.size(), 1)));
    HIPCHK(c, hipMalloc(&c->d_sbuf, sizeof(double) * tot));
    HIPCHK(c, hipMalloc(&c->d_srhs, sizeof(double) * nv));
    HIPCHK(c, hipMalloc(&c->d_sipiv, sizeof(int32_t) * (nbm + std::max<int64_t>(na, 1))));
    HIPCHK(c, hipMalloc(&c->d_sinfo, sizeof(int32_t) * (nb + 1)));
    HIPCHK(c, hipMalloc(&c->d_sprm, sizeof(int32_t) * (nb + 1) * 2 * m));   // + one for the recursive LU
    HIPCHK(c, copy_sync(c, c->d_sslot, slot.data(), sizeof(int64_t) * slot.size(), hipMemcpyHostToDevice));
    {   // the pattern by rows, for the residuals of the refinement steps
        std::vector<int32_t> rp(nv + 1, 0), ci(std::max<int64_t>(nnz, 1)), pm(std::max<int64_t>(nnz, 1));
        for (int64_t p = 0; p < nnz; ++p) ++rp[rv[p] + 1];
        for (int64_t i = 0; i < nv; ++i) rp[i + 1] += rp[i];
        std::vector<int32_t> next(rp.begin(), rp.end() - 1);
        for (int64_t j = 0; j < nv; ++j)
            for (int32_t p = cp[j]; p < cp[j + 1]; ++p) { const int32_t q = next[rv[p]]++; ci[q] = (int32_t)j; pm[q] = p; }
        HIPCHK(c, hipMalloc(&c->d_srowptr, sizeof(int32_t) * (nv + 1)));
        HIPCHK(c, hipMalloc(&c->d_scolidx, sizeof(int32_t) * ci.size()));
        HIPCHK(c, hipMalloc(&c->d_sperm, sizeof(int32_t) * pm.size()));
        HIPCHK(c, hipMalloc(&c->d_sb, sizeof(double) * nv));
        HIPCHK(c, hipMalloc(&c->d_sx, sizeof(double) * (2 * nv + 1)));   // x | ω | per-row backward errors
        HIPCHK(c, hipMalloc(&c->d_ssr, sizeof(double) * nv));
        HIPCHK(c, hipMalloc(&c->d_ssc, sizeof(double) * nv));
        std::vector<int32_t> cv(std::max<int64_t>(nnz, 1));
        for (int64_t j = 0; j < nv; ++j)
            for (int32_t p = cp[j]; p < cp[j + 1]; ++p) cv[p] = (int32_t)j;
        HIPCHK(c, hipMalloc(&c->d_scol, sizeof(int32_t) * cv.size()));
        HIPCHK(c, copy_sync(c, c->d_scol, cv.data(), sizeof(int32_t) * cv.size(), hipMemcpyHostToDevice));
        HIPCHK(c, copy_sync(c, c->d_srowptr, rp.data(), sizeof(int32_t) * (nv + 1), hipMemcpyHostToDevice));
        HIPCHK(c, copy_sync(c, c->d_scolidx, ci.data(), sizeof(int32_t) * ci.size(), hipMemcpyHostToDevice));
        HIPCHK(c, copy_sync(c, c->d_sperm, pm.data(), sizeof(int32_t) * pm.size(), hipMemcpyHostToDevice));
    }
    if (!pad.empty()) HIPCHK(c, copy_sync(c, c->d_spad, pad.data(), sizeof(int64_t) * pad.size(), hipMemcpyHostToDevice));
    if (!c->blas) BLASCHK(c, rocblas_create_handle(&c->blas));
    BLASCHK(c, rocblas_set_stream(c->blas, s));
    c->sps = per_step; c->snst = nstep; c->sna = na; c->sg = g; c->sm = m; c->snb = nb; c->stot = tot;
    c->snpad = (int64_t)pad.size();
    if (m_out) *m_out = m;
    if (nb_out) *nb_out = nb;
    return MH_OK;
}

// Banded block LU with partial pivoting (LAPACK gbtrf at block level) of the step part T: block
// column k (RecLU; rocSOLVER's getrf beyond kPT·kPR rows) pivots over block rows k and k+1,
// whose row interchanges reach block columns k+1 and k+2 (the fill) and the columns of W as one
// composed permutation. The border is taken by bordering: W ← T⁻¹[B | b], then
// the Schur complement S = E − C T⁻¹B (dense, na × na) gives a = S⁻¹(b_A − C T⁻¹b) and the step
// unknowns are T⁻¹b − T⁻¹B a. The system is equilibrated first (its unknowns Λ, X, U, A differ
// in scale by many orders), and one step of iterative refinement on the residual of the sparse
// system follows unless the componentwise backward error is already within 8 ε.
int mh_big_solve(mh_ctx* c, const double* rhs, double* x) {
    if (!c || !c->d_sbuf || c->sm <= 0 || !c->d_bval || !rhs || !x) return MH_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    rocblas_handle h = c->blas;
    BLASCHK(c, rocblas_set_stream(h, s));
    const int64_t nv = c->bnv, nsp = c->sps * c->snst, m = c->sm, nb = c->snb, na = c->sna, nbm = nb * m;
    const int64_t* o = c->so;
    double* B = c->d_sbuf;
    const int im = (int)m, ina = (int)na, iw = (int)(na + 1), l4 = (int)(4 * m), lw = (int)nbm, inbm = (int)nbm;
    const double one = 1., mone = -1.;
    const rocblas_operation N = rocblas_operation_none;
    hipLaunchKernelGGL(chain::fill_f64, dim3((unsigned)((c->stot + 255) / 256)), dim3(256), 0, s, c->stot, B, 0.);
    hipLaunchKernelGGL(bsolve_rowscale, dim3(grid(nv)), dim3(kBlock), 0, s, nv, c->d_srowptr, c->d_sperm, c->d_bval,
                       c->d_ssr);
    hipLaunchKernelGGL(bsolve_colscale, dim3(grid(nv)), dim3(kBlock), 0, s, nv, c->d_bcolptr, c->d_browval, c->d_bval,
                       c->d_ssr, c->d_ssc);
    if (c->bnnz) hipLaunchKernelGGL(bsolve_fill, dim3(grid(c->bnnz)), dim3(kBlock), 0, s, c->bnnz, c->d_sslot, c->d_bval,
                                    c->d_browval, c->d_scol, c->d_ssr, c->d_ssc, B);
    if (c->snpad) hipLaunchKernelGGL(bsolve_ones, dim3(grid(c->snpad)), dim3(kBlock), 0, s, c->snpad, c->d_spad, B);
    HIPCHK(c, hipMemcpyAsync(c->d_sb, rhs, sizeof(double) * nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bsolve_rhs, dim3(grid(nv)), dim3(kBlock), 0, s, nv, nsp, nbm, na, o[1], o[3], c->d_sb, c->d_ssr,
                       B);
    hipLaunchKernelGGL(chain::fill_i32, dim3((unsigned)((nb + 1 + 255) / 256)), dim3(256), 0, s, nb + 1, c->d_sinfo, 0);
    if (int r = check_launch(c, "big solve fill")) return r;
    auto col = [&](int64_t j) { return B + o[0] + j * 4 * m * m; };   // block column j, row 0 (block row j−2)
    double* W = B + o[1];
    double* C = B + o[2];
    double* E = B + o[3];
    double* z = W + nbm * na;   // the step part of the right-hand side / solution
    double* a = E + na * na;    // the border part
    // forward elimination of `ncol` columns of W from its column `c0` (the factor sweep does the
    // factorization at the same time, the refinement steps only the last column)
    auto permute = [&](double* A, int lda, int ncol, int rows, const int32_t* prm) {
        hipLaunchKernelGGL(perm_rows, dim3(ncol), dim3(256), sizeof(double) * rows, s, rows, A, lda, prm);
    };
    auto forward = [&](bool factor, int c0, int ncol) -> int {
        for (int64_t k = 0; k < nb; ++k) {
            double* P = col(k) + 2 * m;   // D_k, then L_{k+1} below it
            int32_t* piv = c->d_sipiv + k * m;
            int32_t* prm = c->d_sprm + k * 2 * m;
            const bool two = k + 1 < nb;
            const int Mr = two ? 2 * im : im;
            if (factor) {
                if (Mr <= kPT * kPR) {
                    RecLU lu{c, s, h, l4, c->d_sinfo + k, c->d_sprm + nb * 2 * m};
                    if (int r = lu.run(P, Mr, im, piv, 0)) return r;
                    if (int r = check_launch(c, "block LU")) return r;
                } else {
                    BLASCHK(c, rocsolver_dgetrf(h, Mr, im, P, l4, piv, c->d_sinfo + k));
                }
                hipLaunchKernelGGL(piv_to_perm, dim3(1), dim3(256), sizeof(int32_t) * (Mr + im), s, im, Mr, piv, 0, prm);
                for (int jj = 1; jj <= 2; ++jj) {
                    if (k + jj >= nb) break;
                    double* T = col(k + jj) + (2 - jj) * m;   // block rows k, k+1 of column k+jj
                    permute(T, l4, im, Mr, prm);
                    BLASCHK(c, rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, N, rocblas_diagonal_unit, im, im,
                                             &one, P, l4, T, l4));
                    BLASCHK(c, rocblas_dgemm(h, N, N, im, im, im, &mone, P + m, l4, T, l4, &one, T + m, l4));
                }
            }
            double* Wk = W + (int64_t)c0 * nbm + k * m;
            permute(Wk, lw, ncol, Mr, prm);
            BLASCHK(c, rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, N, rocblas_diagonal_unit, im, ncol, &one, P,
                                     l4, Wk, lw));
            if (two) BLASCHK(c, rocblas_dgemm(h, N, N, im, ncol, im, &mone, P + m, l4, Wk, lw, &one, Wk + m, lw));
        }
        return MH_OK;
    };
    // X_k = U_kk⁻¹ (W_k − U_{k,k+1} X_{k+1} − U_{k,k+2} X_{k+2}) over the same columns
    auto backward = [&](int c0, int ncol) -> int {
        for (int64_t k = nb - 1; k >= 0; --k) {
            double* Xk = W + (int64_t)c0 * nbm + k * m;
            for (int jj = 1; jj <= 2; ++jj) {
                if (k + jj >= nb) break;
                BLASCHK(c, rocblas_dgemm(h, N, N, im, ncol, im, &mone, col(k + jj) + (2 - jj) * m, l4, Xk + jj * m, lw,
                                         &one, Xk, lw));
            }
            BLASCHK(c, rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, N, rocblas_diagonal_non_unit, im, ncol,
                                     &one, col(k) + 2 * m, l4, Xk, lw));
        }
        return MH_OK;
    };
    if (int r = forward(true, 0, iw)) return r;
    if (int r = backward(0, iw)) return r;
    if (na) {
        BLASCHK(c, rocblas_dgemm(h, N, N, ina, iw, inbm, &mone, C, ina, W, lw, &one, E, ina));   // [S | b_A − C z]
        BLASCHK(c, rocsolver_dgetrf(h, ina, ina, E, ina, c->d_sipiv + nbm, c->d_sinfo + nb));
        BLASCHK(c, rocsolver_dgetrs(h, N, ina, 1, E, ina, c->d_sipiv + nbm, a, ina));
        BLASCHK(c, rocblas_dgemv(h, N, inbm, ina, &mone, W, lw, a, 1, &one, z, 1));
    } else {
        HIPCHK(c, hipMemsetAsync(c->d_sinfo + nb, 0, sizeof(int32_t), s));
    }
    hipLaunchKernelGGL(bsolve_sol, dim3(grid(nv)), dim3(kBlock), 0, s, nv, nsp, nbm, na, o[1], o[3], B, c->d_ssc, c->d_sx);
    for (int it = 0; it < 1; ++it) {
        hipLaunchKernelGGL(bsolve_resid, dim3(grid(nv)), dim3(kBlock), 0, s, nv, c->d_srowptr, c->d_scolidx,
                           c->d_sperm, c->d_bval, c->d_sx, c->d_sb, c->d_srhs, c->d_sx + nv + 1);
        // refine only if x is not already a componentwise backward-stable solution
        hipLaunchKernelGGL(bsolve_omega, dim3(1), dim3(kBlock), 0, s, nv, c->d_sx + nv + 1, c->d_sx + nv);
        double omega = 0.;
        HIPCHK(c, hipMemcpyAsync(&omega, c->d_sx + nv, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        if (omega <= 8. * DBL_EPSILON) break;
        hipLaunchKernelGGL(bsolve_rhs, dim3(grid(nv)), dim3(kBlock), 0, s, nv, nsp, nbm, na, o[1], o[3], c->d_srhs,
                           c->d_ssr, B);
        if (int r = forward(false, ina, 1)) return r;
        if (int r = backward(ina, 1)) return r;
        if (na) {
            BLASCHK(c, rocblas_dgemv(h, N, ina, inbm, &mone, C, ina, z, 1, &one, a, 1));
            BLASCHK(c, rocsolver_dgetrs(h, N, ina, 1, E, ina, c->d_sipiv + nbm, a, ina));
            BLASCHK(c, rocblas_dgemv(h, N, inbm, ina, &mone, W, lw, a, 1, &one, z, 1));
        }
        hipLaunchKernelGGL(bsolve_sol, dim3(grid(nv)), dim3(kBlock), 0, s, nv, nsp, nbm, na, o[1], o[3], B, c->d_ssc,
                           c->d_srhs);
        hipLaunchKernelGGL(bsolve_add, dim3(grid(nv)), dim3(kBlock), 0, s, nv, c->d_srhs, c->d_sx);
    }
    if (int r = check_launch(c, "big solve gather")) return r;
    std::vector<int32_t> info(nb + 1);
    HIPCHK(c, hipMemcpyAsync(x, c->d_sx, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(info.data(), c->d_sinfo, sizeof(int32_t) * (nb + 1), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    for (int64_t k = 0; k <= nb; ++k)
        if (info[k] > 0) {
            c->err = "zero pivot " + std::to_string(info[k]) + (k < nb ? " in block column " + std::to_string(k + 1)
                                                                        : std::string(" in the border block"));
            return MH_ERR_SOLVE;
        }
    for (int64_t i = 0; i < nv; ++i)
        if (!std::isfinite(x[i])) { c->err = "non-finite solution"; return MH_ERR_SOLVE; }
    return MH_OK;
}

}  // extern "C"
