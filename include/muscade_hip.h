/* muscade_hip.h — C ABI of the MI355X element-evaluation-and-assembly library.
 *
 * The library replaces the per-element loop of Muscade.jl's assembler
 * (src/Assemble.jl:470-505 `assemble!`/`assemble_!`, with the SweepX `addin!`
 * methods of src/SweepX.jl:39-96 and the scatter helpers `add_value!`/`add_∂!`
 * of src/Assemble.jl:559-608) and the one-off index-map construction of
 * `asmvec!`/`asmmat!` (src/Assemble.jl:340-448) that `prepare` runs
 * (src/SweepX.jl:24-33).
 *
 * Conventions
 *   - Every function returns an int status: 0 = MH_OK, otherwise an MH_ERR_*
 *     code; mh_last_error() gives the message. Nothing throws across the ABI.
 *   - Host pointers are borrowed for the duration of the call only.
 *   - Dof indices crossing the ABI are 1-based int64 like Muscade's `asm`
 *     and `Disassembler` arrays; 0 means "not in the dof group" (skip).
 *   - Arrays described as "n×nele" are column-major with the n entries of
 *     one element contiguous (Julia's Matrix layout, e.g. asm[1,ieletyp]).
 *   - A context owns one HIP stream and all device buffers; it is not
 *     re-entrant (one host thread per context).
 */
#ifndef MUSCADE_HIP_H
#define MUSCADE_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    MH_OK = 0,
    MH_ERR_ARG = 1,     /* invalid argument or call order                      */
    MH_ERR_HIP = 2,     /* HIP runtime error                                   */
    MH_ERR_NAN = 3,     /* an element returned NaN (Assemble.jl:625-631)       */
    MH_ERR_ORIENT = 4,  /* orient2 nearly parallel to a beam (BeamElement.jl:135) */
    MH_ERR_SIZE = 5,    /* model exceeds the 32-bit index range of the device maps */
    MH_ERR_SOLVE = 6    /* linear solve failed: zero or non-finite pivot (SweepX.jl:205-208) */
};

/* Element kinds with a device kernel. Each Muscade element type (one
 * `ieletyp`, i.e. one entry of model.eleobj) becomes one group of one kind. */
enum {
    MH_KIND_BEAM = 1,    /* EulerBeam3D{BeamCrossSection,false} (toolbox/BeamElement.jl:87-174)        */
    MH_KIND_BEAM_U = 2,  /* EulerBeam3D{BeamCrossSection,true}: 3 U-dofs, load per length             */
    MH_KIND_HOLD = 3,    /* Hold = DofConstraint{:X} equal, gap=x (src/BasicElements.jl:426-438,484-488) */
    MH_KIND_DOFLOAD = 4, /* DofLoad, R = -value(t) evaluated by the host (BasicElements.jl:275-284)      */
    MH_KIND_ZERO = 5,    /* Lagrangian-only element (DofCost, BasicElements.jl:32-50): in SweepX its
                            X-dof pairs enter the pattern and contribute structural zeros             */
    MH_KIND_BAR = 6,     /* Bar3D{Mat,false} with AxisymmetricBarCrossSection (toolbox/BarElement.jl:89-202) */
    MH_KIND_BAR_U = 7,   /* Bar3D{Mat,true}: 3 U-dofs, load per length (BarElement.jl:115-118,192)     */
    MH_KIND_HOST = 8     /* residual evaluated by the host with user functions (DofConstraint{:X} with a
                            gap Functor, BasicElements.jl:426-438; user elements that define
                            `residual`, ElementAPI.jl:144-173): the host hands R and ∂R/∂δX to
                            mh_set_host_blocks (and ∂R/∂δr to mh_set_host_dr) before each assembly;
                            the device scatters them                                                  */
};

typedef struct mh_ctx mh_ctx;

/* Error detail of the last assembly: the first element (in Muscade's
 * traversal order) whose residual or tangent contained a NaN. */
typedef struct mh_err {
    int32_t code;
    int32_t igroup; /* 0-based group (= ieletyp-1) */
    int64_t iele;   /* 0-based element within the group (= iele-1) */
} mh_err;

/* ---- context ------------------------------------------------------------ */
int mh_ctx_create(int device, mh_ctx** ctx);
int mh_ctx_destroy(mh_ctx* ctx);
const char* mh_last_error(const mh_ctx* ctx);
const char* mh_strerror(int code);

/* ---- model description (Model + Disassembler, src/ModelDescription.jl,
 *      src/Assemble.jl:32-99) ------------------------------------------------
 * Start a model: dof counts per class and node coordinates xyz (3×nnod). */
int mh_model_begin(mh_ctx* ctx, int64_t nXdof, int64_t nUdof, int64_t nAdof, int64_t nnod, const double* xyz);
/* Add one element group, in Muscade's eletype order.
 *   idxX   nX×nele  model X-dof of each element X-dof (dis.dis[ieletyp].index[iele].X)
 *   scaleX nX       dis.dis[ieletyp].scale.X
 *   idxU   nU×nele  model U-dof of each element U-dof (MH_KIND_BEAM_U, MH_KIND_BAR_U, else NULL)
 *   nodes  2×nele   1-based node numbers of each beam or bar (else NULL)
 *   orient2 3×nele  first bending axis of each beam (EulerBeam3D's orient2 keyword,
 *                   or column 2 of its rₘ; beams only, NULL for bars)
 *   mat    16×nele  beams: BeamCrossSection fields in declaration order
 *          10×nele  bars: AxisymmetricBarCrossSection fields in declaration order, then ϵₛ
 *                   (Bar3D's keyword, BarElement.jl:106; eps() by default)
 * Returns the 0-based group number in *igroup. */
int mh_group_add(mh_ctx* ctx, int kind, int64_t nele, int nX, const int64_t* idxX, const double* scaleX, int nU,
                 const int64_t* idxU, const int64_t* nodes, const double* orient2, const double* mat, int* igroup);

/* EulerBeam3DA: the stiffnesses of mask bits k (1 EA, 2 EI₂, 4 EI₃, 8 GJ) of beam group igroup
 * become stiffness·10^a, a = A[idxA[e]] (idxA: 1×nele, 1-based model A-dofs). A wrapper in the
 * style of the reference's Spring, which scales its L₀ and EA by exp10 of its A-dofs
 * (test/SomeElements.jl:166-186), so that A-dofs parametrize the beam's section
 * (BeamElement.jl:151-174 with the scaled BeamCrossSection). Every kernel then evaluates the
 * scaled section at the device A (mh_set_A); mh_vjp_lagrangian adds ∂L/∂A. Call before prepare. */
int mh_group_set_adofs(mh_ctx* ctx, int igroup, const int64_t* idxA, int mask);
/* ∂R/∂a of every element of such a group at the device X (mh_set_X, ider 0) and A: 12 numbers
 * per element, element-major, unscaled; ∂R/∂a = ln 10 · R(X; the masked stiffnesses alone,
 * no loads), R being linear in the scaled stiffnesses. The A column of the residual path's
 * Jacobian in a DirectXUA step (L2[Λ,A], src/DirectXUA.jl:85-120 with an element whose R
 * reads A, as the reference's test Spring does). MH_ERR_NAN names the first NaN element. */
int mh_beam_dRdA(mh_ctx* ctx, int igroup, double* out);

/* ---- prepare(AssemblySweepX{OX}) (src/SweepX.jl:24-33) ------------------
 * Builds on the device the CSC pattern of Lλx and the sorted
 * contribution→nzval maps that replace asm[1,:] / asm[2,:]
 * (asmvec!/asmmat!, src/Assemble.jl:340-448). OX (0, 1, 2) is the order of
 * the assembler, as in the type AssemblySweepX{OX}; mh_assemble_sweepx must be
 * called with the same OX. Returns nnz. */
int mh_prepare_sweepx(mh_ctx* ctx, int OX, int64_t* nnz);
/* The pattern exactly as Muscade's SparseMatrixCSC: colptr (nX+1), rowval (nnz), 1-based. */
int mh_get_pattern(mh_ctx* ctx, int64_t* colptr, int64_t* rowval);
/* asm[1,ieletyp] (nX×nele) and asm[2,ieletyp] (nX²×nele, entry i+nX·(j-1)), 1-based. */
int mh_get_asm(mh_ctx* ctx, int igroup, int64_t* asm1, int64_t* asm2);

/* ---- state (State, src/Assemble.jl:106-159) -----------------------------
 * Upload X[ider] (ider 0,1,2 = X, X′, X″), U[0] and A. Device-resident until replaced. */
int mh_set_X(mh_ctx* ctx, int ider, const double* x);
int mh_set_U(mh_ctx* ctx, const double* u);
int mh_set_A(mh_ctx* ctx, const double* a);
/* Host-evaluated element values (DofLoad: value(t,args...) per element). */
int mh_set_point_values(mh_ctx* ctx, int igroup, const double* values);
/* state.time, passed to the residuals as t (Assemble.jl:484; Bar3D's weight ramp,
 * BarElement.jl:144). */
int mh_set_time(mh_ctx* ctx, double t);

/* ---- adjoint (vector-Jacobian product) of the residual Lagrangian ----------
 * L(Λ,X,U,A) = Σ_e Λ_eᵀ R_e(X,U,A,t): the Lagrangian Muscade forms from residual
 * elements (getlagrangian: L = Λ∘₁R, src/Assemble.jl:707-736; DirectXUA's residual
 * path, src/DirectXUA.jl:96-120). After mh_set_lambda (Λ, nX values) and the state
 * (mh_set_X/U/A, mh_set_time), mh_vjp_lagrangian forms ∂L/∂X⁽ᵈᵉʳ⁾ (der = 0..OX) and
 * ∂L/∂U on the device, in unscaled model coordinates, recomputing each element's
 * internal state (tangent, time jets) instead of storing it; the per-element products
 * are reduced into the model vectors in Muscade's element order (no atomics).
 * mh_get_vjp copies them out: gX is (OX+1)×nX (der-major), gU nU, gA nA. gA gathers
 * ∂L/∂a of the EulerBeam3DA groups (mh_group_set_adofs); it is zero without them (EulerBeam3D,
 * Bar3D, Hold and DofLoad have no A-dofs: BeamElement.jl:109-118, BarElement.jl:107-114,
 * BasicElements.jl:275-284,406-438). There is no
 * counterpart in the reference (which forms such gradients with forward-mode duals,
 * src/Diagnostic.jl:851-911); it serves the gradient (rrule) of a solve's Lagrangian. */
int mh_set_lambda(mh_ctx* ctx, const double* lam);
int mh_vjp_lagrangian(mh_ctx* ctx, int OX, mh_err* err);
int mh_get_vjp(mh_ctx* ctx, double* gX, double* gU, double* gA);

/* ---- assemble!{mission}(::AssemblySweepX{OX}) (src/Assemble.jl:470-505,
 *      src/SweepX.jl:39-96) ---------------------------------------------------
 * mission: 0 = :step, 1 = :iter. c6 = Newmark coefficients (a₁,a₂,a₃,b₁,b₂,b₃),
 * Newmarkβcoefficients{OX} (src/SweepX.jl:3-15). Results stay on the device
 * (out.Lλ and out.Lλx.nzval); fetch them with mh_get_rhs / mh_get_nzval. */
int mh_assemble_sweepx(mh_ctx* ctx, int OX, int mission, const double* c6, mh_err* err);
/* The same assembly launched as one captured HIP graph (no counterpart in the reference;
 * a launch-overhead option for the Newton iterations of SweepX.jl:186-221). The graph is
 * captured on the first call and replayed while OX, the dof scales and, when a bar group is
 * present (its kernel takes the time), the time (mh_set_time) are unchanged (one graph is kept
 * per mission, :step and :iter); any change recaptures it. c6 and the state setters
 * (mh_set_X/U/A, mh_set_point_values) act through device buffers and need no recapture: without
 * bars, a whole time integration replays two graphs. Results are bitwise those of mh_assemble_sweepx;
 * mh_last_timings then gives the whole graph's time in ms[0] (ms[1] = ms[2] = 0). */
int mh_assemble_sweepx_graph(mh_ctx* ctx, int OX, int mission, const double* c6, mh_err* err);
/* Number of graph captures so far (a measurement utility). */
int mh_graph_captures(mh_ctx* ctx, int64_t* n);
int mh_get_rhs(mh_ctx* ctx, double* rhs);
int mh_get_nzval(mh_ctx* ctx, double* nzval);
/* Device pointers of the assembled arrays (valid until the next prepare). */
int mh_device_outputs(mh_ctx* ctx, double** d_rhs, double** d_nzval);
/* Per-element results of the last assembly for one group (for element-level
 * checks, like diffed_residual): R (nX×nele) and ∂R/∂X·scale (nX²×nele). */
int mh_get_element_blocks(mh_ctx* ctx, int igroup, double* R, double* K);

/* Element results of a MH_KIND_HOST group for the next SweepX assemblies: R (nX×nele) and
 * ∂R/∂δX (nX²×nele, entry i+nX·j of an element's block), unscaled; the library applies the
 * group's dof scales (R_i·s_i, K_ij·s_i·s_j, the convention of the device kinds). For a static
 * element (a DofConstraint) ∂R/∂δX = ∂R/∂X and its ∂R/∂δr (:step) is zero. */
int mh_set_host_blocks(mh_ctx* ctx, int igroup, const double* R, const double* K);
/* ∂R/∂δr of a MH_KIND_HOST group's elements for the next :step assembly (nX×nele, unscaled; the
 * library applies scale.X): the rhs term −C·a − M·b that `add_∂!{1,:minus}(out.Lλ,asm[1],iele,
 * Lλ,1:Nx,(Nx+1,))` subtracts in SweepX{1,2}'s :step (src/SweepX.jl:45-58,70-80), for host
 * elements with inertia or damping. K of mh_set_host_blocks is then ∂R/∂δX along the trial path
 * (K + a₁C + b₁M, src/SweepX.jl:53-66). mh_set_host_blocks resets ∂R/∂δr to zero. */
int mh_set_host_dr(mh_ctx* ctx, int igroup, const double* dR);

/* ---- element results: getresult / @espy of EulerBeam3D (src/Output.jl:131-181,
 *      toolbox/BeamElement.jl:28-64,151-174) ---------------------------------------
 * The requestables of every beam of group igroup at the device state (mh_set_X/U, X up to
 * ND time levels, ND = 1..3), with no host re-evaluation: per element
 *   ε[ND], κ[ND][3]                     (☼ε, ♢κ = (vₗ₂[1], vₗ₂[3], −vₗ₂[2])·2/L)
 *   per Gauss point (4): x[ND][3], κgp[ND][3], fᵢ, mᵢ[3], fₑ[3], mₑ[3]
 *                                      (☼x, ☼κgp; the resultants ☼fᵢ, ☼mᵢ, ☼fₑ as returned by
 *                                       `resultants`, before the U-load, and ☼mₑ)
 * that is nres = 4·ND + 4·(6·ND + 10) numbers, followed by ngauge strain-gauge strains
 * (StrainGaugeOnEulerBeam3D's ☼ε, toolbox/StrainGaugeOnBeamElement.jl:62-76) for the gauge
 * tables `gauges` (4×ngauge×nele: E, K1, K2, K3 of each gauge of each element, as the
 * wrapper's constructor computes them from P and D, StrainGaugeOnBeamElement.jl:57-66).
 * out: (nres+ngauge)×nele, one element's numbers contiguous. */
int mh_beam_results(mh_ctx* ctx, int igroup, int ND, int ngauge, const double* gauges, double* out);

/* Derivatives of the strain-gauge requestables for costs on them (ElementCost wrapping a
 * StrainGaugeOnEulerBeam3D, src/BasicElements.jl:117-132, in DirectXUA's Lagrangian path,
 * src/DirectXUA.jl:121-171): for the elements sel (1-based, nsel of them; NULL = all) of
 * beam group igroup at the device X, the values of εₐₓ, κ₁, κ₂, κ₃, their gradients (4×12)
 * and their Hessians (4×78, upper triangle row-major) w.r.t. the 12 element X-dofs in
 * unscaled model coordinates, exactly as revariate{2} (src/Taylor.jl:143-151) forms them.
 * out: 364×n. The gauge strains are linear in these (E, K1, K2, K3). */
int mh_beam_gauge_derivs(mh_ctx* ctx, int igroup, int64_t nsel, const int64_t* sel, double* out);
/* mh_beam_gauge_derivs that allocates nothing. It has no counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments and results. The selection and the
 * output staging live in a fixed device array of 11.9 MB that each GPU the library runs on
 * holds until the process ends, so a call takes at most 4096 elements (MH_ERR_SIZE beyond).
 * Every call on one GPU shares that array: calls that overlap on one GPU (contexts driven
 * from several host threads) race. */
int mh_beam_gauge_derivs_noalloc(mh_ctx* ctx, int igroup, int64_t nsel, const int64_t* sel, double* out);

/* ---- AssemblyDirect{OX,OU,IA}: one DirectXUA step (src/DirectXUA.jl:18-150) ----
 * prepare(AssemblyDirect{OX,OU,IA},model,dis) (DirectXUA.jl:22-56) builds the class-pair
 * patterns over the classes Λ, X, U (Λ and X share the X dofs): nnz4 = nnz of the patterns
 * (Λ|X, Λ|X), (Λ|X, U), (U, Λ|X), (U, U), fetched 1-based with mh_get_direct_pattern
 * (which = 0..3). Beams and bars contribute by the residual path (DirectXUA.jl:85-120):
 * L1[Λ] += R·Δt, L2[Λ,β][1,j] += ∂R/∂β⁽ʲ⁾·scale.β·Δt, L2[β,Λ][j,1] += its transpose; Hold and
 * DofLoad by the Lagrangian path of L = Λ∘₁R (:121-171), which needs Λ (mh_set_lambda).
 * mission: 0 = :vectors (L1 only), 1 = :matrices. mh_get_direct returns L1 =
 * [Λ (nX) | X (OX+1)·nX | U (OU+1)·nU] and L2 = the pairs (α,β) ∈ {Λ,X,U}² in row-major order
 * without (Λ,Λ), each nder(α)×nder(β) blocks of nnz(α,β) values (nder = 1, OX+1, OU+1).
 * Elements with user cost functions (DofCost) are not evaluated on the device. A-dofs enter only
 * through A-scaled beams (mh_group_set_adofs): their section follows the device A (mh_set_A),
 * and their A column is mh_beam_dRdA's, outside these blocks. */
int mh_set_group_scales(mh_ctx* ctx, int igroup, const double* scaleL, const double* scaleU);
int mh_prepare_direct(mh_ctx* ctx, int OX, int OU, int IA, int64_t* nnz4);
int mh_get_direct_pattern(mh_ctx* ctx, int which, int64_t* colptr, int64_t* rowval);
int mh_assemble_direct(mh_ctx* ctx, int mission, double dt, mh_err* err);
int mh_get_direct(mh_ctx* ctx, double* L1, double* L2);

/* ---- linear solve and Newton decrement on the device, chain topologies ------
 * Replaces SweepX's host factorization and solve, `lu(out.Lλx)` / `lu!` and
 * `Δx = Lλx\out.Lλ` (src/SweepX.jl:205-209), for models whose multi-node elements join
 * their nodes into paths (beam and bar chains, many chains batched in one model), and
 * Newmarkβdecrement!{OX} with the convergence norms (src/SweepX.jl:98-132,210-211), so
 * that a Newton iteration leaves X on the device and returns two scalars. The host
 * (muscade_amd/chainsolve.py) finds the layout once, after prepare:
 *   B           3 or 6: positions per node block (a node's X dofs, padded)
 *   nblock      number of node blocks, in path order (paths one after another)
 *   pos         nX: for X dof d+1, its 1-based position (block·B + slot + 1) in the block
 *               system, or −p for the λ dof of λ pair p (a Hold's Lagrange-multiplier dof,
 *               coupled to its held dof only: src/BasicElements.jl:426-438)
 *   pair_hpos   npair: 1-based position of each pair's held dof
 *   dummy_pos   ndummy: 1-based padding positions (identity rows)
 * Every position must be taken exactly once. mh_chain_solve solves Lλx·Δx = Lλ from the
 * last assembly by block cyclic reduction (partial pivoting inside the B×B blocks);
 * Δx stays on the device (mh_get_dx copies it). MH_ERR_SOLVE: zero or non-finite pivot,
 * as SweepX's "matrix factorization failed". mh_chain_timings: ms of [pack, elimination,
 * back substitution] of the last solve.
 * mh_newton_prepare uploads the dof scales (dis.scaleX); mh_newton_update applies
 * Newmarkβdecrement!{OX} (first = the :step iteration; c6 as for the assembly) to the
 * device X, X′, X″ with the last Δx, in the host's operation order, and returns
 * norms = [Σ Δx², Σ Lλ²]; mh_get_X copies a device X⁽ⁱᵈᵉʳ⁾ out. */
int mh_chain_prepare(mh_ctx* ctx, int B, int64_t nblock, const int64_t* pos, int64_t npair, const int64_t* pair_hpos,
                     int64_t ndummy, const int64_t* dummy_pos);
int mh_chain_solve(mh_ctx* ctx);
int mh_chain_timings(mh_ctx* ctx, double* ms);
int mh_get_dx(mh_ctx* ctx, double* dx);
int mh_newton_prepare(mh_ctx* ctx, const double* scaleX);
int mh_newton_update(mh_ctx* ctx, int OX, int first, const double* c6, double* norms);
int mh_get_X(mh_ctx* ctx, int ider, double* x);
/* One whole device Newton iteration of SweepX{OX} (assembly with mission :step when first and
 * OX > 0, else :iter; the chain solve; the Newmark decrement and the norms) as one captured HIP
 * graph, replayed while OX, first, the buffers and (with bars) the time are unchanged; c6 is
 * read from device memory, so without bars a time integration replays two graphs, one for the
 * first iteration of every step and one for the others (no counterpart in the reference; the loop of
 * SweepX.jl:186-221 with its lu/solve on the device). One host wait per iteration. Errors as
 * mh_assemble_sweepx (MH_ERR_NAN) and mh_chain_solve (MH_ERR_SOLVE); after either, the device X
 * already holds that iteration's update. mh_last_timings: the whole iteration. The assembly
 * scatters Lλx straight into the chain solve's blocks: out.Lλx (mh_get_nzval) is not formed;
 * Lλ (mh_get_rhs) is. */
int mh_newton_iteration_graph(mh_ctx* ctx, int OX, int first, const double* c6, double* norms, mh_err* err);
/* The adjoint solve of a limit state g(X) at an equilibrium (no counterpart in the reference;
 * the step before the adjoint gradient of a reliability or design-point search): with the
 * tangent of the last assembly (SweepX{0} :iter, symmetric), solve K·μ = rhs (rhs = ∂g/∂X, nX
 * values in unscaled model coordinates) by the chain solve and set Λ := alpha·μ on the device,
 * ready for mh_vjp_lagrangian (alpha = −1 gives dg/dA = Λᵀ∂R/∂A, dg/dU = Λᵀ∂R/∂U, since
 * R(X*(A,U),A,U) = 0). Needs mh_chain_prepare and mh_newton_prepare (the dof scales). */
int mh_adjoint_solve(mh_ctx* ctx, const double* rhs, double alpha);
/* The same with ∂g/∂X given by its n nonzeros (idx: 1-based model X-dofs, repeated dofs
 * summed; val: their values): a limit state reads few dofs, and nothing of size nX crosses PCIe. */
int mh_adjoint_solve_sparse(mh_ctx* ctx, int64_t n, const int64_t* idx, const double* val, double alpha);

/* ---- DirectXUA's all-steps system on the device (src/DirectXUA.jl:245-383:
 *      makepattern / preparebig / assemblebig!, with SparseTools.prepare / addin!,
 *      src/SparseTools.jl:32-137, and the finite-difference weights of
 *      src/FiniteDifferences.jl:2-34) --------------------------------------------------
 * After mh_prepare_direct on a context whose model holds every step (copies of the model,
 * one per step), mh_big_prepare takes the np contributions to the nv×nv system once: for
 * contribution p, its 0-based row and column, its source src (an index into the context's
 * per-step block values — L2[Λ,X][1,d] (d = 0..OX), L2[X,Λ][d,1], L2[Λ,U][1,1], L2[U,Λ][1,1]
 * in the layout of mh_get_direct's blocks, each nnz of its pattern — or, from
 * 2(OX+1)·nnz(XX) + 2·nnz(XU) on, an index into the nhost host-provided values) and its weight
 * w (the stencil weights and Δt powers). It sorts them into a CSC pattern (nnz returned,
 * mh_big_get_pattern: 1-based) and the contribution maps. mh_big_assemble, after each
 * mh_assemble_direct, sums each nz's contributions on the device in contribution order and
 * copies the values out (nzval = NULL: the values stay on the device). */
int mh_big_prepare(mh_ctx* ctx, int64_t nv, int64_t np, const int64_t* rows, const int64_t* cols, const int64_t* src,
                   const double* w, int64_t nhost, int64_t* nnz);
int mh_big_get_pattern(mh_ctx* ctx, int64_t* colptr, int64_t* rowval);
int mh_big_assemble(mh_ctx* ctx, const double* hostvals, double* nzval);

/* The LU and back substitution of the all-steps system on the device (in place of
 * `lu(Lvv)`, `Δv = Lvv\Lv` on the host, src/DirectXUA.jl:489-490). The unknowns are
 * nstep steps of per_step unknowns each (Λ, X, U of one step), then na border unknowns (the
 * A-dofs), so that nv = per_step·nstep + na. Finite differences couple steps at most a few
 * apart: mh_big_solve_prepare reads that width g off the pattern and lays the matrix out as
 * dense blocks of g steps (order m = g·per_step, returned with their count nb), block
 * tridiagonal and bordered by the A rows and columns. MH_ERR_SIZE when m > 4000 (the row
 * permutations pass through LDS) or the blocks do not fit in the device memory. mh_big_solve factors the values of the last mh_big_assemble (which
 * may then pass nzval = NULL: nothing leaves the device) by block LU, partial pivoting
 * inside the diagonal blocks, and returns x (nv) for the right-hand side rhs (nv).
 * MH_ERR_SOLVE: a zero pivot or a non-finite solution. */
int mh_big_solve_prepare(mh_ctx* ctx, int64_t per_step, int64_t nstep, int64_t na, int64_t* m, int64_t* nb);
int mh_big_solve(mh_ctx* ctx, const double* rhs, double* x);

/* ---- measurement -----------------------------------------------------------
 * HIP-event times (ms) of the last assembly, measured on the context's stream:
 * ms[0] element evaluation (all groups), ms[1] nzval scatter, ms[2] rhs scatter. */
int mh_last_timings(mh_ctx* ctx, double* ms);
/* Measurement utility (no counterpart in the reference): one streaming copy of
 * n doubles on the context's stream, time in *ms. It is the known-size kernel
 * that HBM byte counters (FETCH_SIZE, WRITE_SIZE) are calibrated against. */
int mh_calib_copy(mh_ctx* ctx, int64_t n, double* ms);
/* Shape of the nz scatter picked by the prepare-time tuner (0..15: segments per lane ×
 * LDS chunk, with or without nontemporal index loads and stores; see mh_lib.hip
 * launch_scatter) and each shape's measured time in ms (16).
 * A measurement utility; no counterpart in the reference. */
int mh_get_tuning(mh_ctx* ctx, int* variant, float* ms);
/* The static beam kernel's LDS tile (elements per workgroup: 64, 128 or 256; every tile gives
 * the same numbers), picked by a tuner at the context's first SweepX assembly on its state, and
 * each tile's measured time in ms (3); MH_ERR_ARG before that assembly. A measurement utility;
 * no counterpart in the reference. */
int mh_get_beam_tuning(mh_ctx* ctx, int* block, float* ms);
/* Override the tuner's pick (a testing and measurement utility: every shape gives bitwise the
 * same sums). */
int mh_set_scatter_variant(mh_ctx* ctx, int variant);
int mh_sync(mh_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
