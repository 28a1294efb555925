/*
 * sem_abi.h -- C ABI of libsemhip.so, the MI355X (gfx950) operator-apply-and-
 * solve path of SpectralElements.jl.
 *
 * Conventions
 *  - Every nodal field is a device array of n_local = (nr*Ex)*(ns*Ey) values in
 *    the reference's local layout: column-major (nr*Ex, ns*Ey), element
 *    (ex,ey) owning rows ex*nr..ex*nr+nr-1 and columns ey*ns..ey*ns+ns-1
 *    (mesh.jl:94-100). Julia arrays and Fortran-ordered NumPy/torch views pass
 *    through unchanged.
 *  - Small dense matrices (D, interpolation, ABu operands) are column-major.
 *  - Field pointers are device pointers owned by the caller and borrowed for
 *    the duration of the call; host pointers are named *_host.
 *  - All calls are ordered on the context's HIP stream. Calls that return host
 *    scalars synchronise that stream.
 *  - Return value: SEM_OK (0), SEM_WARN_MAXITER (1, the reference's
 *    warn-and-return of pcg.jl:39) or a negative error code; the message of
 *    the last error on the calling thread is sem_last_error().
 */
#ifndef SEM_ABI_H
#define SEM_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEM_ABI_VERSION 1

enum sem_status {
  SEM_OK = 0,
  SEM_WARN_MAXITER = 1,
  SEM_WARN_STALLED = 2, /* sem_pcg_mixed: stopped after four replacements
                           without a 10% gain (see there) */
  SEM_ERR_ARG = -1,
  SEM_ERR_HIP = -2,
  SEM_ERR_UNSUPPORTED = -3,
  SEM_ERR_NONFINITE = -4,
  SEM_ERR_BREAKDOWN = -5
};

/* Operator precision. The MFMA forms (nr == ns == 16) keep FP32 fields and
 * run the Dr/Ds contractions on matrix cores with FP16 or BF16 inputs and
 * FP32 accumulation; the X3 forms split each input into hi + lo and sum
 * three products (lo*hi + hi*lo + hi*hi), about FP32 accuracy. The
 * single-pass forms round each contracted vector after removing its linear
 * interpolant through the two end nodes (added back exactly in FP32), so
 * smooth fields lose no more than the FP32 operator does. */
enum sem_precision {
  SEM_F64 = 0,
  SEM_F32 = 1,
  SEM_F16_MFMA = 2,
  SEM_BF16_MFMA = 3,
  SEM_F16X3_MFMA = 4,
  SEM_BF16X3_MFMA = 5
};

enum sem_deform {
  SEM_DEFORM_NONE = 0,     /* fixU, mesh.jl:6-8 */
  SEM_DEFORM_SINSIN = 1,   /* x += a sin(pi x) sin(pi y), y += a sin(pi x) sin(pi y) */
  SEM_DEFORM_AFFINE01 = 2  /* usrgeom of p2d_explicit.jl:78: [-1,1]^2 -> [0,1]^2 */
};

/* mesh field selectors for sem_mesh_field (names of struct Mesh, mesh.jl:25-64) */
enum sem_mesh_field_id {
  SEM_X = 0, SEM_Y, SEM_JAC, SEM_JACI, SEM_RX, SEM_RY, SEM_SX, SEM_SY,
  SEM_B, SEM_BI, SEM_G11, SEM_G12, SEM_G22, SEM_MULT, SEM_NUM_MESH_FIELDS
};

/* Dirichlet mask bits (generateMask, mesh.jl:149-175: bc = [xmin,xmax,ymin,ymax]) */
enum sem_bc_bits { SEM_BC_XMIN = 1, SEM_BC_XMAX = 2, SEM_BC_YMIN = 4, SEM_BC_YMAX = 8 };

enum sem_precond {
  SEM_PRECOND_IDENTITY = 0, /* opPrecond of Diffusion, diffusion.jl:47-49 */
  SEM_PRECOND_DIAG = 1,     /* z = d .* r with a diagonal d that is equal on all
                               copies of a shared node: the Jacobi inverse
                               (sem_operator_jacobi) or any assembled field */
  SEM_PRECOND_DIAG_LOCAL = 2 /* z = d .* r with a d that may differ between the
                               copies of a node, as 1 ./ (B * b0) with the
                               element-local B of convectionDiffusion.jl:87-91:
                               z is then discontinuous, so <p, A p> is formed on
                               the assembled A p as pcg.jl:50 does (the fused
                               solver's element-local pairing needs a continuous
                               p). sem_pcg only, SEM_F64 operators. */
};

typedef struct sem_ctx sem_ctx;
typedef struct sem_mesh sem_mesh;
typedef struct sem_operator sem_operator;
typedef struct sem_gsmap sem_gsmap;

/* ---------------------------------------------------------------- context */
int sem_abi_version(void);
const char* sem_last_error(void);
/* stream: a hipStream_t (NULL = the legacy default stream) */
int sem_ctx_create(int device, void* stream, sem_ctx** out);
int sem_ctx_destroy(sem_ctx* ctx);
int sem_ctx_set_stream(sem_ctx* ctx, void* stream);
int sem_ctx_synchronize(sem_ctx* ctx);

/* ------------------------------------------------ L0 host primitives (FP64) */
/* FastGaussQuadrature.gausslobatto(n) as called at mesh.jl:70-71, semmesh.jl:11 */
int sem_gausslobatto(int n, double* z_host, double* w_host);
/* derivMat(x), derivMat.jl:9-35; D_host is n*n column-major */
int sem_derivmat(int n, const double* x_host, double* D_host);
/* interpMat(xo, xi), interp.jl:10-35; J_host is no*ni column-major */
int sem_interpmat(int no, const double* xo_host, int ni, const double* xi_host, double* J_host);
/* bdfExtK(t, k), time.jl:31-53: the BDFk/EXTk coefficients of the time
 * history t[0..nt) = [t_{n+1}, t_n, ...]: b = first row of derivMat(unique(t)),
 * a = interpMat(t_{n+1}, unique(t)[1:]), zero-padded to k (a) and k + 1 (b)
 * while the history holds fewer distinct times, truncated otherwise; a single
 * distinct time gives a = [1, 0, ...]. a_host has k entries, b_host k + 1. */
int sem_bdfextk(int nt, const double* t_host, int k, double* a_host, double* b_host);
/* semmesh(E, n), semmesh.jl:9-27: z_host, w_host have E*n entries */
int sem_semmesh(int E, int n, double* z_host, double* w_host);

/* ------------------------------------------------------------------- mesh */
/* Mesh(nr, ns, Ex, Ey, ifperiodic, deform), mesh.jl:66-133. GLL nodes and D
 * on the host; coordinates, the deformation, jac (jac.jl:24-40), B, Bi, G11,
 * G12, G22 and mult are computed in FP64 on the GPU. */
int sem_mesh_create(sem_ctx* ctx, int nr, int ns, int Ex, int Ey, int periodic_x,
                    int periodic_y, int deform_kind, double deform_param, sem_mesh** out);
/* Replaces the coordinates by caller-deformed ones (a user deform callback,
 * mesh.jl:108) and recomputes the geometric factors on the GPU. */
int sem_mesh_set_coords(sem_ctx* ctx, sem_mesh* mesh, const double* x, const double* y);
int sem_mesh_destroy(sem_mesh* mesh);
/* sizes: nr, ns, Ex, Ey, n_local, n_global (any pointer may be NULL) */
int sem_mesh_sizes(const sem_mesh* mesh, int* nr, int* ns, int* Ex, int* Ey,
                   int64_t* n_local, int64_t* n_global);
/* device pointer to a mesh array (read-only, n_local doubles) */
int sem_mesh_field(const sem_mesh* mesh, int which, const double** dev_ptr);
/* device pointers to Dr, Ds (nr*nr, ns*ns column-major) and the GLL nodes
 * and weights zr, wr, zs, ws */
int sem_mesh_dmat(const sem_mesh* mesh, const double** Dr, const double** Ds);
int sem_mesh_gll(const sem_mesh* mesh, const double** zr, const double** wr,
                 const double** zs, const double** ws);
/* local-to-global numbering (int32 device array, 0-based, column-major
 * global grid): the local l2g intended by mesh.jl:87-89 */
int sem_mesh_l2g(const sem_mesh* mesh, const int32_t** dev_l2g);
/* generateMask(bc, msh), mesh.jl:149-175: bc = 4 chars 'D'/'N' for
 * [xmin,xmax,ymin,ymax]; returns SEM_BC_* bits (periodic directions cleared) */
int sem_mask_bits(const sem_mesh* mesh, const char* bc4, int* bits);
/* materialises the mask M (float64 0/1 field) for the given bits */
int sem_mask_fill(sem_ctx* ctx, const sem_mesh* mesh, int bits, double* M);

/* ----------------------------------------------------- gather-scatter map */
/* Deterministic QQ' over a precomputed map: local indices sorted by global
 * id (stable, so copies of a node are in local-index order), segment offsets.
 * Built from any 0-based local->global table (the NNlib form of
 * gatherScatter.jl:27-34 takes such a table). */
int sem_gsmap_create(sem_ctx* ctx, int64_t n_local, const int32_t* l2g_host,
                     int64_t n_global, sem_gsmap** out);
int sem_gsmap_destroy(sem_gsmap* map);
int sem_gsmap_apply(sem_ctx* ctx, const sem_gsmap* map, const double* u, double* y);
/* gatherScatter(u, msh), gatherScatter.jl:18-21 (dense QQtx*u*QQty' in the
 * reference; value-identical here, x-pairs summed before y-pairs) */
int sem_gather_scatter(sem_ctx* ctx, const sem_mesh* mesh, const double* u, double* y);
int sem_gather_scatter_f32(sem_ctx* ctx, const sem_mesh* mesh, const float* u, float* y);

/* ---------------------------------------------------- pointwise and ABu */
/* mask(u, M) = M .* u, mask.jl:10-18 (M == NULL: copy) */
int sem_mask(sem_ctx* ctx, int64_t n, const double* M, const double* u, double* y);
/* mask by bits (structured M = mx my') */
int sem_mask_bc(sem_ctx* ctx, const sem_mesh* mesh, int bits, const double* u, double* y);
/* mass(u, msh) = B .* u, mass.jl:12-22 */
int sem_mass(sem_ctx* ctx, const sem_mesh* mesh, const double* u, double* y);
/* ABu(As, Br, u), ABu.jl:9-37, u is m x n column-major. As (ma x na) or Br
 * (mb x nb) may be NULL for the reference's []. The output is
 * (m*mb/nb) x (n*ma/na). The pullback ABu(As', Br', dv) of ABu.jl:39-41 is
 * ABu with trans = 1 (both operands transposed). */
int sem_abu(sem_ctx* ctx, int m, int n, const double* u, const double* As, int ma,
            int na, const double* Br, int mb, int nb, int trans, double* y);

/* elementwise y = a .* b, a ./ b or alpha ./ a over n entries (y may alias a
 * or b): the products, quotients and inverse of the diagonal operator of the
 * Spectral submodule (DiagonalOp, Spectral/Operators.jl:37-118) */
enum sem_pointwise_op {
  SEM_PW_MUL = 0,        /* y = a .* b */
  SEM_PW_DIV = 1,        /* y = a ./ b */
  SEM_PW_SCALAR_DIV = 2, /* y = alpha ./ a */
  SEM_PW_DIV_SCALAR = 3, /* y = a ./ alpha */
  SEM_PW_SCALE = 4       /* y = alpha .* a */
};
int sem_pointwise(sem_ctx* ctx, int op, int64_t n, double alpha, const double* a,
                  const double* b, double* y);

/* ------------------------------------------------------ element operators */
/* lapl(u, msh), lapl.jl:26-36 (element-local weak Laplacian, no QQ') */
int sem_lapl(sem_ctx* ctx, const sem_mesh* mesh, const double* u, double* y);
/* hlmz(u, nu, k, msh) = nu .* lapl(u) + k .* mass(u), hlmz.jl:12-19; nu_f or
 * k_f (device fields) replace the scalars when non-NULL */
int sem_hlmz(sem_ctx* ctx, const sem_mesh* mesh, const double* u, double nu,
             const double* nu_f, double k, const double* k_f, double* y);

/* wr = G11 .* a + G12 .* b, ws = G12 .* a + G22 .* b over n entries: the
 * pointwise step of laplace (lapl.jl:75-76) and of its dealiased form on the
 * interpolation grid (lapl.jl:95-96) */
int sem_gprod(sem_ctx* ctx, int64_t n, const double* G11, const double* G12, const double* G22,
              const double* a, const double* b, double* wr, double* ws);
/* out = alpha .* x .+ beta .* y over n entries (the broadcast sums of
 * lapl.jl:78 and hlmz.jl:15-16); out may alias x or y */
int sem_axpby(sem_ctx* ctx, int64_t n, double alpha, const double* x, double beta,
              const double* y, double* out);

/* ------------------------------------------------------- fused operator */
typedef struct sem_op_desc {
  double nu;             /* scalar nu (used when nu_f == NULL) */
  double k;              /* scalar mass coefficient (bdfB[1] in diffusion.jl:40) */
  const double* nu_f;    /* optional nu field, post-multiplies A u (hlmz.jl:15) */
  const double* k_f;     /* optional k field */
  const double* G11;     /* optional geometric-factor overrides (the 12-argument */
  const double* G12;     /*   lapl of lapl.jl:54-68 takes G11/G12/G22 from the   */
  const double* G22;     /*   caller, e.g. nu folded in: nu .* G, semPS.jl:161)  */
  const double* B;       /* optional mass override */
  int bc_bits;           /* Dirichlet mask bits (0 = no mask) */
  int assemble;          /* 1: y = M .* QQ'(nu A u + k B u) (opLHS, diffusion.jl:36-45);
                            0: element-local nu A u + k B u (hlmz.jl:12) */
} sem_op_desc;

/* Builds a device operator. precision SEM_F64 keeps FP64 factors; SEM_F32
 * stores them in FP32; SEM_F16_MFMA / SEM_BF16_MFMA (nr == ns == 16) run the
 * Dr/Ds contractions as FP16/BF16-input, FP32-accumulate MFMA with FP32
 * factors and fields (SEM_F16X3_MFMA / SEM_BF16X3_MFMA: split inputs, three
 * products). */
int sem_operator_create(sem_ctx* ctx, const sem_mesh* mesh, const sem_op_desc* desc,
                        int precision, sem_operator** out);
int sem_operator_destroy(sem_operator* op);
int sem_operator_precision(const sem_operator* op, int* precision);
/* y = op(u); u and y are double* for SEM_F64 and float* otherwise */
int sem_operator_apply(sem_ctx* ctx, const sem_operator* op, const void* u, void* y);
/* Jacobi inverse diagonal dinv = M .* 1 ./ QQ'(diag_local), FP64 */
int sem_operator_jacobi(sem_ctx* ctx, const sem_operator* op, double* dinv);

/* --------------------------------------------------------------- solvers */
typedef struct sem_pcg_info {
  int iters;        /* iterations performed (k of pcg.jl) */
  int converged;    /* 1 if norm(r, Inf) <= tol */
  double res;       /* final norm(r, Inf) of the recursive residual */
  double elapsed_ms;/* device time of the solve */
  int outer;        /* defect-correction outer iterations (mixed solver) */
  int inner_total;  /* total inner iterations (mixed solver) */
} sem_pcg_info;

/* pcg!(x, b, opA; opM, mult, tol, maxiter), pcg.jl:16-79, fused on the GPU:
 * x0 = 0 (pcg.jl:25; pcg! ignores the incoming x, :71), stop when
 * norm(r, Inf) <= tol (absolute, :36), mult-weighted inner products with the
 * operator mesh's mult (:45, :52), SEM_WARN_MAXITER at k == maxiter (:39).
 * tol < 0 selects the relative test norm(r, Inf) <= |tol| * norm(b, Inf),
 * with norm(b, Inf) formed on the device (no host read before the solve;
 * time steppers use it). precond: SEM_PRECOND_IDENTITY, SEM_PRECOND_DIAG or
 * SEM_PRECOND_DIAG_LOCAL with the device diagonal d (z = d .* r). For SEM_F32
 * operators b, x and d are float*. x must not alias b. Returns once the stop
 * test has fired (info filled in); the last update of x may still be queued
 * on ctx's stream, so readers of x on other streams synchronise with it. */
int sem_pcg(sem_ctx* ctx, const sem_operator* op, int precond, const void* d,
            const void* b, void* x, double tol, int maxiter, sem_pcg_info* info);

/* sem_pcg without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device or pinned host memory during a solve
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels, results and status codes as sem_pcg;
 * its workspace (r, p, w, reduction partials, device scalars) is a fixed
 * device array of 24 MiB + 128 KiB per GPU, held until the process ends, and
 * the device state is read back by plain copies. It fits n_local up to about
 * 1,048,000 FP64 (2,096,000 FP32) local nodes; larger meshes return
 * SEM_ERR_UNSUPPORTED before any launch. Every call on one GPU shares that
 * array, so calls that overlap on one GPU (several streams or threads) race;
 * each call clears what must start at zero. */
int sem_pcg_noalloc(sem_ctx* ctx, const sem_operator* op, int precond, const void* d,
                    const void* b, void* x, double tol, int maxiter, sem_pcg_info* info);

/* Mixed precision: FP64 defect correction by residual replacement ("reliable
 * updates"). One fused PCG runs with the low-precision operator op_lo on FP32
 * vectors; its iterate is accumulated into x in FP64 and its recursive
 * residual is replaced by the true residual b - op64(x) (FP64 operator)
 * whenever it has dropped by the factor inner_tol since the last replacement
 * (when the search direction is kept, also after 1000 inner iterations). The search direction and iteration count
 * carry over (the Krylov space is not restarted); at each replacement the
 * inner vectors are rescaled by a power of two so the residual is O(1). A
 * cycle whose true residual is not finite or grew tenfold is rolled back and
 * the Krylov space restarted. FP32, FP16 and split (x3) operators keep the
 * search direction across replacements until one shows no progress, then
 * restart it at every replacement; the single-pass BF16 operator always restarts
 * it (classical defect correction).
 * SEM_MIXED_RELIABLE=1 / 0 forces one policy. Stops when norm(b - op64(x), Inf) <= tol;
 * SEM_WARN_MAXITER after max_outer replacements or max_inner total inner
 * iterations; SEM_WARN_STALLED after four replacements without a 10%
 * improvement (the low-precision operator's error times the condition number
 * is then too large for defect correction to contract). Either way x holds
 * the best iterate and info->res its true residual. d64: FP64 Jacobi
 * diagonal (or NULL for identity). info->outer = replacements, info->iters =
 * inner iterations. Environment SEM_DEBUG_MIXED=1 prints one line per cycle
 * to stderr. */
int sem_pcg_mixed(sem_ctx* ctx, const sem_operator* op64, const sem_operator* op_lo,
                  const double* d64, const double* b, double* x, double tol,
                  double inner_tol, int max_outer, int max_inner, sem_pcg_info* info);

/* --------------------------------------------- multigrid preconditioner */
/* A stronger opM for pcg than the identity / Jacobi of diffusion.jl:47-49
 * (the reference's own next step is the commented element fast-diagonalisation
 * sketch of lapl.jl:105-118 and p2d_explicit.jl:108-141). A symmetric V-cycle:
 * p-multigrid on the element grid (N -> N/2 -> ... -> 1), then h-multigrid on
 * bilinear elements (2 x 2 elements -> 1 while both element counts are even
 * and at least 4), each level the fused operator mask(QQ'(nu A + k B))
 * rediscretised on a mesh whose nodes sample the fine geometry, Chebyshev-
 * Jacobi smoothing of degree `degree` on [lmax/ratio, lmax] (lmax = safety x
 * the Rayleigh quotient after npow power iterations from a hashed start
 * vector), Chebyshev of degree `degree_coarse` on the coarsest level.
 * Prolongation interpolates per element; restriction is its transpose
 * (QQ'_c(J' (mult .* r) J), masked). Needs an FP64, assembled operator with
 * scalar nu and k (no field or factor overrides). */
typedef struct sem_mg_params {
  int degree;         /* default 3 */
  double ratio;       /* default 20 */
  int degree_coarse;  /* default 8 */
  int npow;           /* default 20 */
  double safety;      /* default 1.1 */
  int max_levels;     /* default 32 */
  int precision;      /* SEM_F64 (default) or SEM_F32: the V-cycle (smoothing,
                         transfers, level operators) in FP32, with the FP64
                         eigenvalue bounds; sem_pcg_mg then uses the flexible
                         (Polak-Ribiere) beta, the rest of the solve in FP64 */
} sem_mg_params;

typedef struct sem_mg sem_mg;

/* fills the defaults listed above */
int sem_mg_default_params(sem_mg_params* params);
/* builds the hierarchy for op (params NULL: defaults); the operator and its
 * mesh must outlive the object */
int sem_mg_create(sem_ctx* ctx, const sem_operator* op, const sem_mg_params* params,
                  sem_mg** out);
int sem_mg_destroy(sem_mg* mg);
int sem_mg_num_levels(const sem_mg* mg, int* nlevels);
/* level 0 is the operator's own mesh; p = points per direction */
int sem_mg_level_info(const sem_mg* mg, int level, int* p, int* Ex, int* Ey, double* lmax);
/* replaces a level's lmax (the smoothing interval follows) */
int sem_mg_set_lmax(sem_mg* mg, int level, double lmax);
/* z = V(r): one V-cycle applied to an assembled, masked residual r (FP64) */
int sem_mg_vcycle(sem_ctx* ctx, sem_mg* mg, const double* r, double* z);
/* pcg(b, opA; opM = V-cycle, mult, tol, maxiter), pcg.jl:16-60: x0 = 0,
 * stop when norm(r, Inf) <= tol (absolute), mult-weighted inner products,
 * SEM_WARN_MAXITER at k == maxiter. op must be the operator mg was built for. */
int sem_pcg_mg(sem_ctx* ctx, const sem_operator* op, sem_mg* mg, const double* b, double* x,
               double tol, int maxiter, sem_pcg_info* info);

/* ------------------------------------ mesh generation, point evaluation */
/* gordonHall(xrm,xrp,xsm,xsp,yrm,yrp,ysm,ysp,zr,zs), geom.jl:8-31:
 * transfinite interpolation of four boundary curves onto the nr x ns grid of
 * the points r (nr) x s (ns) in [-1,1]: x = (1-r)/2 xrm(s) + (1+r)/2 xrp(s)
 * + (1-s)/2 xsm(r) + (1+s)/2 xsp(r) - the bilinear interpolant of the
 * corners (taken from the r-edges). xrm, xrp, yrm, yrp have ns entries (the
 * r = -1 and r = +1 edges, indexed by s), xsm, xsp, ysm, ysp nr entries.
 * All arrays are device arrays; x, y are nr x ns column-major. With r, s the
 * 1-D node coordinates of a whole mesh (semmesh) this maps the mesh. The
 * reference's corner term transposes the corner matrix (geom.jl:16-23), which
 * differs from the transfinite interpolant unless x(1,-1) = x(-1,1); this
 * is the interpolant the construction intends. */
int sem_gordon_hall(sem_ctx* ctx, int nr, int ns, const double* r, const double* s,
                    const double* xrm, const double* xrp, const double* xsm, const double* xsp,
                    const double* yrm, const double* yrp, const double* ysm, const double* ysp,
                    double* x, double* y);
/* semInterp(u, xp, yp, rxe, rye, sxe, sye, zr, zs, xe, ye, tol, maxiter),
 * interp.jl:47-102, for npts points at once over the Ex x Ey elements of an
 * (nr x ns)-node layout: each point is located by scanning the elements
 * whose bounding box holds it and running Newton on the element map from the
 * closest GLL node (r += rx dx + ry dy, s += sx dx + sy dy with the inverse
 * Jacobian interpolated at (r, s); the reference's update reads dx twice and
 * dy twice, interp.jl:80-81) until |(dx, dy)| < tol within maxiter steps
 * and inside [-1,1]^2; out = u interpolated there, NaN if no element
 * converges; elem (optional) = the element (ex + Ex ey) or -1. Fields are
 * device arrays in the local layout; xp, yp, out, elem are device arrays. */
int sem_interp_points(sem_ctx* ctx, int nr, int ns, int Ex, int Ey, const double* u,
                      const double* x, const double* y, const double* rx, const double* ry,
                      const double* sx, const double* sy, int npts, const double* xp,
                      const double* yp, double tol, int maxiter, double* out, int* elem);
/* the same over a mesh (its x, y, rx, ry, sx, sy) */
int sem_mesh_interp(sem_ctx* ctx, const sem_mesh* mesh, const double* u, int npts,
                    const double* xp, const double* yp, double tol, int maxiter, double* out,
                    int* elem);

/* ----------------------------------------------------- vector utilities */
/* sum(a .* b .* mult) with the mesh's mult (pcg.jl:45); result to host */
int sem_wdot(sem_ctx* ctx, const sem_mesh* mesh, const double* a, const double* b,
             double* result_host);
/* sum(a .* b .* w) over n entries (w NULL: sum(a .* b); b NULL: sum(a .* w)
 * or sum(a)); result to host.
 * The local-node sums of the objectives, e.g. 1/2 sum(B .* e .* e)
 * (semPS.jl:210-211) */
int sem_dot(sem_ctx* ctx, int64_t n, const double* a, const double* b, const double* w,
            double* result_host);
/* acc_dev[0] += scale * sum(a .* b .* w) with acc_dev in device memory: the
 * same sums accumulated over a time loop without a host read per step (the
 * objective J = sum_n dt/2 sum(B .* u^n .* u^n) of the time-dependent
 * adjoints, semPS.jl:210-211 per step) */
int sem_dot_acc(sem_ctx* ctx, int64_t n, const double* a, const double* b, const double* w,
                double scale, double* acc_dev);
/* y = [QQ'] mask(sum_j coef[j] (B .* if weighted[j]) x[j]): the composition of
 * makeRHS! (convectionDiffusion.jl:93-110, diffusion.jl:51-65): mass terms,
 * explicit history terms, then mask and gatherScatter. nterms <= 8; x[j] are
 * device pointers in a host array; bc_bits < 0 skips the mask; assemble = 0
 * skips QQ'. y may alias any x[j]. */
int sem_combine(sem_ctx* ctx, const sem_mesh* mesh, int nterms, const double* coef,
                const double* const* x, const int* weighted, int bc_bits, int assemble,
                double* y);
/* norm(a, Inf); result to host */
int sem_norm_inf(sem_ctx* ctx, int64_t n, const double* a, double* result_host);

/* ------------------------------------------------------ adjoint kernels */
/* grad(u, msh), grad.jl:15-34: ux = rx ur + sx us, uy = ry ur + sy us */
int sem_grad(sem_ctx* ctx, const sem_mesh* mesh, const double* u, double* ux, double* uy);
/* adjoint of grad as a map L -> 2L: Dr'(rx vx + ry vy) + (sx vx + sy vy) Ds
 * (the transpose of grad.jl:15-34; gradT of grad.jl:44-63 split per input) */
int sem_grad_t(sem_ctx* ctx, const sem_mesh* mesh, const double* vx, const double* vy,
               double* y);
/* d/dnu of lambda' A(nu) u for A(nu) = lapl with G_ij <- nu .* G_ij
 * (semPS.jl:161-165): g = B .* (grad(lambda) . grad(u)) per local node, the
 * per-element kernel of the adjoint gradient (adjoint.jl:13-17) */
int sem_adjoint_dnu(sem_ctx* ctx, const sem_mesh* mesh, const double* u,
                    const double* lam, double* g);
/* G_ij .* nu for a variable-coefficient diffusion operator (semPS.jl:161-163,
 * without the dealiasing interpolation): pass the outputs as the G overrides
 * of sem_op_desc */
int sem_vardiff_factors(sem_ctx* ctx, const sem_mesh* mesh, const double* nu, double* G11,
                        double* G12, double* G22);
/* advect(T, vx, vy, mshV, mshD, Jr, Js), advect.jl:45-64: dealiased
 * Js'Jr'[B_D (J vx . J grad T)] with Jr = interpMat(zr_D, zr_V) */
int sem_advect(sem_ctx* ctx, const sem_mesh* meshV, const sem_mesh* meshD,
               const double* T, const double* vx, const double* vy, double* y);
/* transpose of advect with respect to T (for the discrete adjoint) */
int sem_advect_t(sem_ctx* ctx, const sem_mesh* meshV, const sem_mesh* meshD,
                 const double* w, const double* vx, const double* vy, double* y);

/* ------------------------------------- convection-diffusion time loop */
/* The step loop of ConvectionDiffusion (convectionDiffusion.jl:5-179) run by
 * the library: sem_cd_steps repeats step!(cdn) (:148-155) nsteps times with
 * f = 0 and ub = 0 (homogeneous Dirichlet data on the bc_bits sides):
 * updateHist! of the field and the TimeStepper (mesh.jl:199-215,
 * time.jl:101-112), bdfExtK (time.jl:31-53), makeRHS! (:93-110, advect of a
 * history computed once and shifted with it), solve! (:112-122) as the fused
 * Jacobi PCG of mask(QQ'(hlmz(., nu, bdfB[1]))) with the relative stop test
 * norm(r, Inf) <= tol * norm(rhs, Inf). One operator and Jacobi diagonal per
 * distinct bdfB[1] is built on first use and kept. nu: scalar, or the field
 * nu_f (device, kept by the caller). maxiter 0: the number of local nodes
 * (pcg.jl:22). Order k in [1, 4]. */
typedef struct sem_cd sem_cd;
int sem_cd_create(sem_ctx* ctx, const sem_mesh* meshV, const sem_mesh* meshD, int bc_bits,
                  int k, double nu, const double* nu_f, const double* vx, const double* vy,
                  double dt, double tol, int maxiter, sem_cd** out);
int sem_cd_destroy(sem_cd* cd);
/* State (caller-owned device fields, host arrays updated in place): u;
 * uh[k] the histories and adv[k] the cached advect(uh_i) buffers (the
 * pointer arrays rotate like the histories), adv_valid[k] whether adv[i]
 * holds advect(uh[i]); t[k+1] the TimeStepper times. Optional outputs:
 * J_dev[0] += dt/2 sum(B .* u .* u) after every step (device); a_out
 * (nsteps x k) and b_out (nsteps x (k+1)) the per-step coefficients;
 * iters_total the PCG iterations summed. SEM_WARN_MAXITER if some solve hit
 * maxiter. The steps issue the same launches as the step-by-step entry
 * points, in the same order (identical results). */
int sem_cd_steps(sem_ctx* ctx, sem_cd* cd, int nsteps, double* u, double** uh, double** adv,
                 int* adv_valid, double* t_host, double* J_dev, double* a_out_host,
                 double* b_out_host, int64_t* iters_total);

#ifdef __cplusplus
}
#endif
#endif /* SEM_ABI_H */
