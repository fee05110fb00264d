// extern "C" entry points of libmomwire_hip.so (declared in include/momwire_hip.h).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "mw_common.h"

hipError_t mw_launch_same_edge(const mw_mesh& M, double k, hipStream_t st);
hipError_t mw_launch_same_edge_static(const mw_mesh& M, double* out, hipStream_t st);
hipError_t mw_launch_fill_far32(const mw_mesh& M, double k, double omega, double eps, double mu,
                                double2* Z, int64_t ldz, const int32_t* pairs, int32_t n_pairs,
                                bool symmetric, double near_ratio, hipStream_t st);
int64_t mw_same_edge_blocks(int32_t n_se);
hipError_t mw_launch_fill(const mw_mesh& M, double k, double omega, double eps, double mu,
                          double2* Z, int64_t ldz, const int32_t* pairs, int32_t n_pairs,
                          bool symmetric, hipStream_t st);
int64_t mw_getrf_ws_bytes(int64_t n);
int64_t mw_sysv_ws_bytes(int64_t n);
template <typename T>
int mw_getrf_t(int64_t n, int64_t nc, void* A, int64_t lda, int32_t* ipiv, void* work,
               int32_t* info, const volatile int32_t* cancel, hipStream_t st, hipError_t* herr);
template <typename T>
int mw_gesv_t(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv, void* work,
              int32_t* info, const volatile int32_t* cancel, hipStream_t st, hipError_t* herr);
template <typename T>
int mw_sysv_t(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv, void* work,
              int32_t* info, const volatile int32_t* cancel, const int32_t* first_row,
              hipStream_t st, hipError_t* herr);
template <typename T>
int mw_getrs_t(int64_t n, int64_t nrhs, const void* A, int64_t lda, const int32_t* ipiv,
               const void* work, void* X, int64_t ldx, hipStream_t st, hipError_t* herr);
template <typename T>
int mw_sytrs_t(int64_t n, int64_t nrhs, const void* A, int64_t lda, const int32_t* ipiv,
               const void* work, void* X, int64_t ldx, hipStream_t st, hipError_t* herr);
hipError_t mw_gemm_sub_z(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                         const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st);
hipError_t mw_gemm_sub_c(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                         const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st);
hipError_t mw_gemm_tn_add_z(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                            const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st);
hipError_t mw_gemm_tn_add_c(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                            const void* B, int64_t ldb, void* C, int64_t ldc, hipStream_t st);
hipError_t mw_launch_sparse_rows(int64_t nrows, int maxnz, const int32_t* idx, const double* val,
                                 const void* X, int64_t ldx, int64_t ncols, void* out,
                                 int64_t ldo, hipStream_t st);
hipError_t mw_launch_scatter_cols(int64_t n, int64_t ncols, int maxnz, const int32_t* idx,
                                  const double* val, void* X, int64_t ldx, hipStream_t st);
hipError_t mw_launch_z2c(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
                         int64_t ldd, hipStream_t st);
hipError_t mw_launch_c2z(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst,
                         int64_t ldd, int accumulate, hipStream_t st);
hipError_t mw_launch_readout_check(int64_t n, int64_t p, const void* X0, int64_t ldx,
                                   const void* C, int64_t ldc, const void* T, int64_t ldt, void* Y,
                                   int64_t ldy, double tol, void* scratch, double* stats,
                                   int32_t* gate, hipStream_t st);
hipError_t mw_launch_zentries(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                              double k, double omega, double eps, double mu, const int32_t* rows,
                              const int32_t* cols, int64_t n, int table, void* out,
                              hipStream_t st);
hipError_t mw_launch_hm_dense(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                              double k, double omega, double eps, double mu,
                              const int32_t* rows_pool, const int32_t* cols_pool,
                              const mw_hblock* blk, const int64_t* items, int64_t n_items,
                              int table, void* store, hipStream_t st);
hipError_t mw_launch_hm_aca(const mw_mesh& M, const int32_t* wing_seg, const int32_t* wing_slot,
                            double k, double omega, double eps, double mu,
                            const int32_t* rows_pool, const int32_t* cols_pool, mw_hblock* blk,
                            const int32_t* list, int32_t n_list, int32_t n_large, double tol,
                            void* store, void* flags, hipStream_t st);
hipError_t mw_launch_hm_apply(const mw_hblock* blk, const void* store, const mw_hm_titem* t_items,
                              int32_t n_t, const int32_t* t_rows, const int32_t* colpos,
                              const mw_hm_ritem* r_items, int32_t n_r, const int32_t* r_ent,
                              const int32_t* sigma, const void* X, int64_t ldx, int32_t s,
                              void* T, void* Y, int64_t ldy, hipStream_t st);
hipError_t mw_launch_hm_invert(const mw_hblock* blk, const int32_t* list, int32_t n_list,
                               void* store, void* work, int32_t* info, hipStream_t st);
void mw_set_pivot_bound(double lbound);
double mw_get_pivot_bound();
hipError_t mw_launch_batch_fill(const mw_sweep_prob* P, const mw_mesh* meshes, int nprob,
                                const int32_t* se_items, int64_t n_se_items,
                                const mw_kgroup* groups, const int32_t* group_prob,
                                const int32_t* item_group, const int32_t* item_pair,
                                int64_t n_items, int64_t n_plain, double eps, double mu,
                                hipStream_t st);
int mw_sweep_lu_run(const mw_sweep_prob* dP, const mw_ports* dports, const mw_sweep_prob* hp,
                    int nprob, int max_ports, int max_con, int64_t max_scatter,
                    int gemm_variant, double lbound, const uint32_t* items,
                    const int64_t* off, const volatile int32_t* cancel, hipStream_t st,
                    hipError_t* herr);
void mw_sweep_work(const mw_sweep_prob* hp, int nprob, int gemm_variant,
                   std::vector<uint32_t>& items, std::vector<int64_t>& off);
int mw_sweep_wg_run(const mw_sweep_prob* dP, const mw_ports* dports, int nprob, int two,
                    int per_cu, double lbound, int32_t* d_next, const volatile int32_t* cancel,
                    hipStream_t st, hipError_t* herr);
hipError_t mw_launch_add_sparse(int64_t nnz, const int32_t* rows, const int32_t* cols,
                                const void* vals, void* Z, int64_t ldz, hipStream_t st);
hipError_t mw_launch_small_solve(int ns, int nt, void* S, int64_t lds, void* T, int64_t ldt,
                                 int32_t* info, hipStream_t st);

hipError_t mw_launch_somm_six(double er, double ei, double k2, const double* rho, const double* h,
                              int64_t n, double rtol, int form, double2* out,
                              const volatile int32_t* cancel, hipStream_t st);
hipError_t mw_launch_somm_surfaces(double er, double ei, double k2, double omega, double mu,
                                   const double* R1, const double* th, int64_t n, double rtol,
                                   double2* out, const volatile int32_t* cancel, hipStream_t st);
hipError_t mw_launch_ground_fields(const mw_somm_grid& G, double gz, double k, int n_pts,
                                   const double* pts, int n_src, const double* mid,
                                   const double2* mom, int64_t ldm, int p, int accumulate,
                                   double2* out, int64_t ldo, hipStream_t st);
hipError_t mw_launch_somm_eval(const mw_somm_grid& G, const double* R1, const double* th, int64_t n,
                               double2* out, hipStream_t st);
hipError_t mw_launch_fill_ground(const mw_mesh& M, double k, double omega, double eps, double mu,
                                 double gz, double c2r, double c2i, const mw_somm_grid* G,
                                 const double* refl, int nqs, double2* Z, int64_t ldz,
                                 const int32_t* pairs, int32_t n_pairs, bool symmetric,
                                 hipStream_t st);

hipError_t mw_launch_somm_surfaces_varied(const double* eps, double k2, double omega, double mu,
                                          const double* R1, const double* th, int64_t n, double rtol,
                                          double2* out, const volatile int32_t* cancel,
                                          hipStream_t st);
hipError_t mw_launch_element_moments(int n_el, int n_node, int p, const int32_t* tap_i,
                                     const double* tap_v, const int32_t* el_tab, const double* dl,
                                     const int32_t* node_tab, const double2* X, int64_t ldx,
                                     double2* mom, int64_t ldm, double2* delta, int64_t ldd,
                                     hipStream_t st);
hipError_t mw_launch_far_moments(int n_el, const double* mid, const double2* mom, int64_t ldm,
                                 int p, int n_theta, int n_phi, const double* trig_t,
                                 const double* trig_p, double k, int ground, double ground_z,
                                 double eps_re, double eps_im, double2* m_theta, double2* m_phi,
                                 int64_t ldo, hipStream_t st);
hipError_t mw_launch_near_fields(int n_pts, const double* pts, int n_el, const double* mid,
                                 const double2* mom, int64_t ldm, int n_node, const double* nodes,
                                 const double2* delta, int64_t ldd, double node_sign, int p,
                                 double k, double radius, double eta, int magnetic,
                                 int accumulate, double2* out, int64_t ldo, hipStream_t st);
int64_t mw_mlp_n_params();
int64_t mw_mlp_n_w16();
int64_t mw_mlp_ws_bytes(int64_t n);
hipError_t mw_mlp_launch_sync(const mw_mlp& net, hipStream_t st);
hipError_t mw_mlp_launch_forward(const mw_mlp& net, const float* x, int64_t n, float* out, void* ws,
                                 hipStream_t st);
hipError_t mw_mlp_launch_backward(const mw_mlp& net, const float* x, const float* t, int64_t n,
                                  float* loss_out, void* ws, hipStream_t st);
hipError_t mw_mlp_launch_adam(const mw_mlp& net, float lr, float b1, float b2, float eps, int step,
                              hipStream_t st);
hipError_t mw_mlp_launch_grid(const mw_mlp& net, double eps_re, double eps_im, const double* R1,
                              const double* th, const int32_t* idx, int64_t n, double2* vals,
                              double sre, double sim, void* ws, hipStream_t st);

namespace {
thread_local char g_err[512] = "";

int fail(int code, const char* fmt, const char* what = "", long long a = 0, long long b = 0) {
  snprintf(g_err, sizeof g_err, fmt, what, a, b);
  return code;
}

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return MW_OK;
  snprintf(g_err, sizeof g_err, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
  return MW_EHIP;
}

int check_mesh(const mw_mesh* M) {
  if (!M) return fail(MW_EINVAL, "%smesh is NULL");
  if (M->degree < 1 || M->degree > 2)
    return fail(MW_EINVAL, "%sdegree %lld not built (1 or 2)", "", M->degree);
  if (M->n_qp < 1 || M->n_qp > 8)
    return fail(MW_EINVAL, "%sn_qp_pair %lld not built (1 .. 8)", "", M->n_qp);
  if (M->n_seg <= 0 || M->n_basis <= 0 || M->n_tile <= 0)
    return fail(MW_EINVAL, "%sempty mesh (n_seg=%lld, n_basis=%lld)", "", M->n_seg, M->n_basis);
  if (!M->seg || !M->seg_edge || !M->edge || !M->edge_i || !M->se_tab || !M->band_poly ||
      !M->bw_loc || !M->bw_slot || !M->tile_b0 || !M->tile_seg || !M->tile_nseg || !M->gl)
    return fail(MW_EINVAL, "%smesh has a NULL table");
  if (M->ek && !M->ek_label) return fail(MW_EINVAL, "%sextended kernel without coaxial labels");
  return MW_OK;
}

int check_mesh_ground(const mw_mesh* M) {
  int rc = check_mesh(M);
  if (rc) return rc;
  if (M->ek && (!M->ek_img_obs || !M->ek_img_src))
    return fail(MW_EINVAL, "%sextended kernel without the image block's coaxial labels");
  return MW_OK;
}

// A cancel flag that a kernel polls must be memory the device can read: host
// memory allocated mapped (mw_cancel_create, hipHostMalloc) or device memory.
// An ordinary (pageable) host address is refused before anything is
// launched; *dev receives the address the kernels dereference.
int device_cancel(const volatile int32_t* cancel, const volatile int32_t** dev) {
  *dev = nullptr;
  if (!cancel) return MW_OK;
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof a);
  const hipError_t e = hipPointerGetAttributes(&a, (const void*)cancel);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // the query's error must not surface at the next call
    return fail(MW_EINVAL, "%scancel flag is not device-accessible memory (allocate it with "
                           "mw_cancel_create)");
  }
  if (a.type == hipMemoryTypeHost && a.devicePointer) {
    *dev = (const volatile int32_t*)a.devicePointer;
    return MW_OK;
  }
  if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) {
    *dev = cancel;
    return MW_OK;
  }
  return fail(MW_EINVAL, "%scancel flag is not device-accessible memory (allocate it with "
                         "mw_cancel_create)");
}

// The address a kernel may read `cancel` through (mapped pinned or device
// memory), or NULL for an ordinary host address; sets no error.
const volatile int32_t* device_view(const volatile int32_t* cancel) {
  if (!cancel) return nullptr;
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof a);
  if (hipPointerGetAttributes(&a, (const void*)cancel) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (a.type == hipMemoryTypeHost && a.devicePointer)
    return (const volatile int32_t*)a.devicePointer;
  if (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) return cancel;
  return nullptr;
}
}  // namespace

extern "C" {

const char* mw_last_error(void) { return g_err; }
int mw_abi_version(void) { return 1; }

int mw_cancel_create(int32_t** flag) {
  g_err[0] = 0;
  if (!flag) return fail(MW_EINVAL, "%sflag is NULL");
  *flag = nullptr;
  void* p = nullptr;
  const hipError_t e = hipHostMalloc(&p, 64, hipHostMallocMapped);
  if (e != hipSuccess) return hip_status(e, "mw_cancel_create");
  *(volatile int32_t*)p = 0;
  *flag = (int32_t*)p;
  return MW_OK;
}

int mw_cancel_destroy(int32_t* flag) {
  g_err[0] = 0;
  if (!flag) return MW_OK;
  return hip_status(hipHostFree(flag), "mw_cancel_destroy");
}

int mw_same_edge_table(const mw_mesh* mesh, double k, void* stream) {
  g_err[0] = 0;
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (!(k > 0.0)) return fail(MW_EINVAL, "%swavenumber must be positive");
  return hip_status(mw_launch_same_edge(*mesh, k, (hipStream_t)stream), "mw_same_edge_table");
}

int mw_same_edge_static(const mw_mesh* mesh, double* out, void* stream) {
  g_err[0] = 0;
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (mesh->n_se > 0 && !out) return fail(MW_EINVAL, "%sout is NULL");
  return hip_status(mw_launch_same_edge_static(*mesh, out, (hipStream_t)stream),
                    "mw_same_edge_static");
}

int mw_fill_z(const mw_mesh* mesh, double k, double omega, double eps, double mu, void* Z,
              int64_t ldz, const int32_t* tile_pairs, int32_t n_pairs, int32_t symmetric,
              void* stream) {
  g_err[0] = 0;
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (!Z) return fail(MW_EINVAL, "%sZ is NULL");
  if (ldz < mesh->n_basis) return fail(MW_EINVAL, "%sldz %lld < n_basis %lld", "", ldz, mesh->n_basis);
  if (symmetric && (!tile_pairs || n_pairs <= 0))
    return fail(MW_EINVAL, "%ssymmetric fill needs the tile-pair list");
  return hip_status(mw_launch_fill(*mesh, k, omega, eps, mu, (double2*)Z, ldz, tile_pairs,
                                   n_pairs, symmetric != 0, (hipStream_t)stream),
                    "mw_fill_z");
}

int mw_fill_z_far32(const mw_mesh* mesh, double k, double omega, double eps, double mu, void* Z,
                    int64_t ldz, const int32_t* tile_pairs, int32_t n_pairs, int32_t symmetric,
                    double near_ratio, void* stream) {
  g_err[0] = 0;
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (!Z) return fail(MW_EINVAL, "%sZ is NULL");
  if (ldz < mesh->n_basis) return fail(MW_EINVAL, "%sldz %lld < n_basis %lld", "", ldz, mesh->n_basis);
  if (symmetric && (!tile_pairs || n_pairs <= 0))
    return fail(MW_EINVAL, "%ssymmetric fill needs the tile-pair list");
  if (!(near_ratio >= 0.0)) return fail(MW_EINVAL, "%snear_ratio must be >= 0");
  return hip_status(mw_launch_fill_far32(*mesh, k, omega, eps, mu, (double2*)Z, ldz, tile_pairs,
                                         n_pairs, symmetric != 0, near_ratio, (hipStream_t)stream),
                    "mw_fill_z_far32");
}

int64_t mw_getrf_workspace_bytes(int64_t n) { return n > 0 ? mw_getrf_ws_bytes(n) : 0; }
int64_t mw_sysv_workspace_bytes(int64_t n) { return n > 0 ? mw_sysv_ws_bytes(n) : 0; }

int mw_set_pivot_threshold(double bound) {
  g_err[0] = 0;
  if (!(bound == bound)) return fail(MW_EINVAL, "%spivot threshold is NaN");
  mw_set_pivot_bound(bound);
  return MW_OK;
}
double mw_get_pivot_threshold(void) { return mw_get_pivot_bound(); }

static int getrf_common(bool dbl, int64_t n, void* A, int64_t lda, int32_t* ipiv, void* work,
                        int32_t* info, const volatile int32_t* cancel, void* stream) {
  g_err[0] = 0;
  if (n <= 0) return fail(MW_EINVAL, "%sn must be positive");
  if (!A || !ipiv || !work || !info) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < n) return fail(MW_EINVAL, "%slda %lld < n %lld", "", lda, n);
  hipError_t herr = hipSuccess;
  int rc = dbl ? mw_getrf_t<double>(n, n, A, lda, ipiv, work, info, cancel, (hipStream_t)stream, &herr)
               : mw_getrf_t<float>(n, n, A, lda, ipiv, work, info, cancel, (hipStream_t)stream, &herr);
  if (rc == MW_EHIP) return hip_status(herr, "getrf");
  if (rc == MW_EABORT) return fail(MW_EABORT, "%sfactorisation cancelled");
  return rc;
}

int mw_zgetrf(int64_t n, void* A, int64_t lda, int32_t* ipiv, void* work, int32_t* info,
              const volatile int32_t* cancel, void* stream) {
  return getrf_common(true, n, A, lda, ipiv, work, info, cancel, stream);
}
int mw_cgetrf(int64_t n, void* A, int64_t lda, int32_t* ipiv, void* work, int32_t* info,
              const volatile int32_t* cancel, void* stream) {
  return getrf_common(false, n, A, lda, ipiv, work, info, cancel, stream);
}

static int gesv_common(bool dbl, int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
                       void* work, int32_t* info, const volatile int32_t* cancel, void* stream) {
  g_err[0] = 0;
  if (n <= 0 || nrhs < 0) return fail(MW_EINVAL, "%sbad sizes n=%lld nrhs=%lld", "", n, nrhs);
  if (!A || !ipiv || !work || !info) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < n + nrhs) return fail(MW_EINVAL, "%slda %lld < n + nrhs %lld", "", lda, n + nrhs);
  hipError_t herr = hipSuccess;
  int rc = dbl ? mw_gesv_t<double>(n, nrhs, A, lda, ipiv, work, info, cancel, (hipStream_t)stream, &herr)
               : mw_gesv_t<float>(n, nrhs, A, lda, ipiv, work, info, cancel, (hipStream_t)stream, &herr);
  if (rc == MW_EHIP) return hip_status(herr, "gesv");
  if (rc == MW_EABORT) return fail(MW_EABORT, "%sfactorisation cancelled");
  return rc;
}

int mw_zgesv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv, void* work,
             int32_t* info, const volatile int32_t* cancel, void* stream) {
  return gesv_common(true, n, nrhs, A, lda, ipiv, work, info, cancel, stream);
}
}  // extern "C"

// ---------------------------------------------------------------------------
// The factor-and-solve with no allocation: its LU workspace (diagonal-block
// inverses, pivot flags, tournament counters and candidates) is a
// file-scope device array instead of a caller-provided buffer.  Each GPU
// holds its own copy of the array from the first call on that GPU until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kScratchBytes = 17LL << 20;  // mw_getrf_workspace_bytes(16384) fits
constexpr int64_t kScratchMaxN = 16384;
constexpr int kScratchDevs = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevs] = {nullptr};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
// The device address of this GPU's scratch array (looked up once per device,
// with that device current).
int scratch_for_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "mw_zgesv_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kScratchDevs)
    return fail(MW_EINVAL, "%sdevice ordinal %lld beyond the scratch table", "", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_status(e, "mw_zgesv_noalloc: hipGetSymbolAddress");
    if (!p) return fail(MW_EHIP, "%shipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MW_OK;
}
}  // namespace

extern "C" {

int mw_zgesv_noalloc(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
                     int32_t* info, const volatile int32_t* cancel, void* stream) {
  g_err[0] = 0;
  if (n <= 0 || nrhs < 0) return fail(MW_EINVAL, "%sbad sizes n=%lld nrhs=%lld", "", n, nrhs);
  if (n > kScratchMaxN || mw_getrf_ws_bytes(n) > kScratchBytes)
    return fail(MW_ERANGE,
                "%sn=%lld is above the static scratch of mw_zgesv_noalloc (17 MiB: "
                "n <= %lld); use mw_zgesv with a workspace",
                "", n, kScratchMaxN);
  void* work = nullptr;
  int rc = scratch_for_current_device(&work);
  if (rc) return rc;
  return gesv_common(true, n, nrhs, A, lda, ipiv, work, info, cancel, stream);
}

int mw_cgesv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv, void* work,
             int32_t* info, const volatile int32_t* cancel, void* stream) {
  return gesv_common(false, n, nrhs, A, lda, ipiv, work, info, cancel, stream);
}

static int sysv_common(bool dbl, int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv,
                       void* work, int32_t* info, const volatile int32_t* cancel,
                       const int32_t* first_row, void* stream) {
  g_err[0] = 0;
  if (n <= 0 || nrhs < 0) return fail(MW_EINVAL, "%sbad sizes n=%lld nrhs=%lld", "", n, nrhs);
  if (!A || !ipiv || !work || !info) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < n + nrhs) return fail(MW_EINVAL, "%slda %lld < n + nrhs %lld", "", lda, n + nrhs);
  hipError_t herr = hipSuccess;
  int rc = dbl ? mw_sysv_t<double>(n, nrhs, A, lda, ipiv, work, info, cancel, first_row,
                                   (hipStream_t)stream, &herr)
               : mw_sysv_t<float>(n, nrhs, A, lda, ipiv, work, info, cancel, first_row,
                                  (hipStream_t)stream, &herr);
  if (rc == MW_EHIP) return hip_status(herr, "sysv");
  if (rc == MW_EABORT) return fail(MW_EABORT, "%sfactorisation cancelled");
  return rc;
}

int mw_zsysv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv, void* work,
             int32_t* info, const volatile int32_t* cancel, const int32_t* first_row,
             void* stream) {
  return sysv_common(true, n, nrhs, A, lda, ipiv, work, info, cancel, first_row, stream);
}
int mw_csysv(int64_t n, int64_t nrhs, void* A, int64_t lda, int32_t* ipiv, void* work,
             int32_t* info, const volatile int32_t* cancel, const int32_t* first_row,
             void* stream) {
  return sysv_common(false, n, nrhs, A, lda, ipiv, work, info, cancel, first_row, stream);
}

static int getrs_common(bool dbl, int64_t n, int64_t nrhs, const void* A, int64_t lda,
                        const int32_t* ipiv, const void* work, void* X, int64_t ldx,
                        void* stream) {
  g_err[0] = 0;
  if (n <= 0 || nrhs < 0) return fail(MW_EINVAL, "%sbad sizes n=%lld nrhs=%lld", "", n, nrhs);
  if (!A || !X || !work) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < n || ldx < nrhs) return fail(MW_EINVAL, "%sbad leading dimension (lda=%lld ldx=%lld)", "", lda, ldx);
  if (nrhs == 0) return MW_OK;
  hipError_t herr = hipSuccess;
  int rc = dbl ? mw_getrs_t<double>(n, nrhs, A, lda, ipiv, work, X, ldx, (hipStream_t)stream, &herr)
               : mw_getrs_t<float>(n, nrhs, A, lda, ipiv, work, X, ldx, (hipStream_t)stream, &herr);
  if (rc == MW_EHIP) return hip_status(herr, "getrs");
  return rc;
}

int mw_zgetrs(int64_t n, int64_t nrhs, const void* A, int64_t lda, const int32_t* ipiv,
              const void* work, void* X, int64_t ldx, void* stream) {
  return getrs_common(true, n, nrhs, A, lda, ipiv, work, X, ldx, stream);
}

static int sytrs_common(bool dbl, int64_t n, int64_t nrhs, const void* A, int64_t lda,
                        const int32_t* ipiv, const void* work, void* X, int64_t ldx,
                        void* stream) {
  g_err[0] = 0;
  if (n <= 0 || nrhs < 0) return fail(MW_EINVAL, "%sbad sizes n=%lld nrhs=%lld", "", n, nrhs);
  if (!A || !X || !work || (n <= 4096 && !ipiv)) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < n || ldx < nrhs) return fail(MW_EINVAL, "%sbad leading dimension (lda=%lld ldx=%lld)", "", lda, ldx);
  if (nrhs == 0) return MW_OK;
  hipError_t herr = hipSuccess;
  int rc = dbl ? mw_sytrs_t<double>(n, nrhs, A, lda, ipiv, work, X, ldx, (hipStream_t)stream, &herr)
               : mw_sytrs_t<float>(n, nrhs, A, lda, ipiv, work, X, ldx, (hipStream_t)stream, &herr);
  if (rc == MW_EHIP) return hip_status(herr, "sytrs");
  return rc;
}

int mw_zsytrs(int64_t n, int64_t nrhs, const void* A, int64_t lda, const int32_t* ipiv,
              const void* work, void* X, int64_t ldx, void* stream) {
  return sytrs_common(true, n, nrhs, A, lda, ipiv, work, X, ldx, stream);
}
int mw_csytrs(int64_t n, int64_t nrhs, const void* A, int64_t lda, const int32_t* ipiv,
              const void* work, void* X, int64_t ldx, void* stream) {
  return sytrs_common(false, n, nrhs, A, lda, ipiv, work, X, ldx, stream);
}
int mw_cgetrs(int64_t n, int64_t nrhs, const void* A, int64_t lda, const int32_t* ipiv,
              const void* work, void* X, int64_t ldx, void* stream) {
  return getrs_common(false, n, nrhs, A, lda, ipiv, work, X, ldx, stream);
}

int mw_zgemm_sub(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda, const void* B,
                 int64_t ldb, void* C, int64_t ldc, void* stream) {
  g_err[0] = 0;
  if (m < 0 || n < 0 || k < 0) return fail(MW_EINVAL, "%snegative size");
  if ((m && n && k) && (!A || !B || !C)) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < k || ldb < n || ldc < n) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_gemm_sub_z(m, n, k, A, lda, B, ldb, C, ldc, (hipStream_t)stream),
                    "mw_zgemm_sub");
}
int mw_cgemm_sub(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda, const void* B,
                 int64_t ldb, void* C, int64_t ldc, void* stream) {
  g_err[0] = 0;
  if (m < 0 || n < 0 || k < 0) return fail(MW_EINVAL, "%snegative size");
  if ((m && n && k) && (!A || !B || !C)) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < k || ldb < n || ldc < n) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_gemm_sub_c(m, n, k, A, lda, B, ldb, C, ldc, (hipStream_t)stream),
                    "mw_cgemm_sub");
}

int mw_zgemm_tn_add(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda, const void* B,
                    int64_t ldb, void* C, int64_t ldc, void* stream) {
  g_err[0] = 0;
  if (m < 0 || n < 0 || k < 0) return fail(MW_EINVAL, "%snegative size");
  if ((m && n && k) && (!A || !B || !C)) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < m || ldb < n || ldc < n) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_gemm_tn_add_z(m, n, k, A, lda, B, ldb, C, ldc, (hipStream_t)stream),
                    "mw_zgemm_tn_add");
}

int mw_cgemm_tn_add(int64_t m, int64_t n, int64_t k, const void* A, int64_t lda, const void* B,
                    int64_t ldb, void* C, int64_t ldc, void* stream) {
  g_err[0] = 0;
  if (m < 0 || n < 0 || k < 0) return fail(MW_EINVAL, "%snegative size");
  if ((m && n && k) && (!A || !B || !C)) return fail(MW_EINVAL, "%sNULL argument");
  if (lda < m || ldb < n || ldc < n) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_gemm_tn_add_c(m, n, k, A, lda, B, ldb, C, ldc, (hipStream_t)stream),
                    "mw_cgemm_tn_add");
}

int mw_zsparse_rows(int64_t nrows, int32_t maxnz, const int32_t* idx, const double* val,
                    const void* X, int64_t ldx, int64_t ncols, void* out, int64_t ldo,
                    void* stream) {
  g_err[0] = 0;
  if (nrows < 0 || ncols < 0 || maxnz < 0) return fail(MW_EINVAL, "%snegative size");
  if (nrows == 0 || ncols == 0) return MW_OK;
  if (!idx || !val || !X || !out) return fail(MW_EINVAL, "%sNULL argument");
  if (ldx < ncols || ldo < ncols) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_sparse_rows(nrows, maxnz, idx, val, X, ldx, ncols, out, ldo,
                                          (hipStream_t)stream),
                    "mw_zsparse_rows");
}

int mw_zscatter_cols(int64_t n, int64_t ncols, int32_t maxnz, const int32_t* idx,
                     const double* val, void* X, int64_t ldx, void* stream) {
  g_err[0] = 0;
  if (n < 0 || ncols < 0 || maxnz < 0) return fail(MW_EINVAL, "%snegative size");
  if (n == 0 || ncols == 0) return MW_OK;
  if (!idx || !val || !X) return fail(MW_EINVAL, "%sNULL argument");
  if (ldx < ncols) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_scatter_cols(n, ncols, maxnz, idx, val, X, ldx, (hipStream_t)stream),
                    "mw_zscatter_cols");
}

int mw_zadd_sparse(int64_t n, int64_t nnz, const int32_t* rows, const int32_t* cols,
                   const void* vals, void* Z, int64_t ldz, void* stream) {
  g_err[0] = 0;
  if (n < 0 || nnz < 0) return fail(MW_EINVAL, "%snegative size");
  if (nnz == 0) return MW_OK;
  if (!rows || !cols || !vals || !Z) return fail(MW_EINVAL, "%sNULL argument");
  if (ldz < n) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_add_sparse(nnz, rows, cols, vals, Z, ldz, (hipStream_t)stream),
                    "mw_zadd_sparse");
}

int mw_element_moments(int32_t n_el, int32_t n_node, int32_t p, const int32_t* tap_i,
                       const double* tap_v, const int32_t* el_tab, const double* dl,
                       const int32_t* node_tab, const void* X, int64_t ldx, void* mom,
                       int64_t ldm, void* delta, int64_t ldd, void* stream) {
  g_err[0] = 0;
  if (n_el < 0 || n_node < 0 || p < 0) return fail(MW_EINVAL, "%snegative size");
  if (p == 0 || n_el + (int64_t)n_node == 0) return MW_OK;
  if (!tap_i || !tap_v || !el_tab || !X) return fail(MW_EINVAL, "%sNULL argument");
  if (n_el && (!dl || !mom)) return fail(MW_EINVAL, "%sNULL element table or moments");
  if (n_node && (!node_tab || !delta)) return fail(MW_EINVAL, "%sNULL node table or steps");
  if (ldx < p || (n_el && ldm < 3 * (int64_t)p) || (n_node && ldd < p))
    return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_element_moments(n_el, n_node, p, tap_i, tap_v, el_tab, dl, node_tab,
                                              (const double2*)X, ldx, (double2*)mom, ldm,
                                              (double2*)delta, ldd, (hipStream_t)stream),
                    "mw_element_moments");
}

int mw_far_moments(int32_t n_el, const double* mid, const void* mom, int64_t ldm, int32_t p,
                   int32_t n_theta, int32_t n_phi, const double* trig_theta,
                   const double* trig_phi, double k, int32_t ground, double ground_z,
                   double eps_re, double eps_im, void* m_theta, void* m_phi, int64_t ldo,
                   void* stream) {
  g_err[0] = 0;
  if (n_el < 0 || p < 0 || n_theta < 0 || n_phi < 0) return fail(MW_EINVAL, "%snegative size");
  if (ground < 0 || ground > 2)
    return fail(MW_EINVAL, "%sground must be 0 (free space), 1 (PEC) or 2 (finite), got %lld", "",
                ground);
  if (!(k > 0.0)) return fail(MW_EINVAL, "%swavenumber must be positive");
  if (p == 0 || (int64_t)n_theta * n_phi == 0) return MW_OK;
  if (!trig_theta || !trig_phi || !m_theta || !m_phi) return fail(MW_EINVAL, "%sNULL argument");
  if (n_el && (!mid || !mom)) return fail(MW_EINVAL, "%sNULL element table or moments");
  if (ldo < p || (n_el && ldm < 3 * (int64_t)p)) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_far_moments(n_el, mid, (const double2*)mom, ldm, p, n_theta, n_phi,
                                          trig_theta, trig_phi, k, ground, ground_z, eps_re,
                                          eps_im, (double2*)m_theta, (double2*)m_phi, ldo,
                                          (hipStream_t)stream),
                    "mw_far_moments");
}

int mw_near_fields(int32_t n_pts, const double* pts, int32_t n_el, const double* mid,
                   const void* mom, int64_t ldm, int32_t n_node, const double* nodes,
                   const void* delta, int64_t ldd, double node_sign, int32_t p, double k,
                   double radius, double eta, int32_t magnetic, int32_t accumulate, void* out,
                   int64_t ldo, void* stream) {
  g_err[0] = 0;
  if (n_pts < 0 || n_el < 0 || n_node < 0 || p < 0) return fail(MW_EINVAL, "%snegative size");
  if (!(k > 0.0)) return fail(MW_EINVAL, "%swavenumber must be positive");
  if (!(radius >= 0.0)) return fail(MW_EINVAL, "%sradius must be >= 0");
  if (n_pts == 0 || p == 0) return MW_OK;
  if (!pts || !out) return fail(MW_EINVAL, "%sNULL argument");
  if (n_el && (!mid || !mom)) return fail(MW_EINVAL, "%sNULL element table or moments");
  if (!magnetic && n_node && (!nodes || !delta)) return fail(MW_EINVAL, "%sNULL nodes or steps");
  if (ldo < 3 * (int64_t)p || (n_el && ldm < 3 * (int64_t)p) || (!magnetic && n_node && ldd < p))
    return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_near_fields(n_pts, pts, n_el, mid, (const double2*)mom, ldm,
                                          magnetic ? 0 : n_node, nodes, (const double2*)delta, ldd,
                                          node_sign, p, k, radius, eta, magnetic, accumulate,
                                          (double2*)out, ldo, (hipStream_t)stream),
                    "mw_near_fields");
}

int mw_zsmall_solve(int32_t ns, int32_t nt, void* S, int64_t lds, void* T, int64_t ldt,
                    int32_t* info, void* stream) {
  g_err[0] = 0;
  if (ns <= 0 || nt <= 0) return MW_OK;
  if (ns > 32 || nt > 64) return fail(MW_ERANGE, "%ssmall solve limited to ns<=32, nt<=64 (got %lld, %lld)", "", ns, nt);
  if (!S || !T) return fail(MW_EINVAL, "%sNULL argument");
  if (lds < ns || ldt < nt) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_small_solve(ns, nt, S, lds, T, ldt, info, (hipStream_t)stream),
                    "mw_zsmall_solve");
}

int mw_z2c(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst, int64_t ldd,
           void* stream) {
  g_err[0] = 0;
  if (rows < 0 || cols < 0) return fail(MW_EINVAL, "%snegative size");
  if (rows && cols && (!src || !dst)) return fail(MW_EINVAL, "%sNULL argument");
  if (lds < cols || ldd < cols) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_z2c(rows, cols, src, lds, dst, ldd, (hipStream_t)stream), "mw_z2c");
}

int mw_c2z(int64_t rows, int64_t cols, const void* src, int64_t lds, void* dst, int64_t ldd,
           int32_t accumulate, void* stream) {
  g_err[0] = 0;
  if (rows < 0 || cols < 0) return fail(MW_EINVAL, "%snegative size");
  if (rows && cols && (!src || !dst)) return fail(MW_EINVAL, "%sNULL argument");
  if (lds < cols || ldd < cols) return fail(MW_EINVAL, "%sbad leading dimension");
  return hip_status(mw_launch_c2z(rows, cols, src, lds, dst, ldd, accumulate, (hipStream_t)stream),
                    "mw_c2z");
}

int mw_zreadout_check(int64_t n, int64_t p, const void* X0, int64_t ldx, const void* C,
                      int64_t ldc, const void* T, int64_t ldt, void* Y, int64_t ldy, double tol,
                      void* scratch, double* stats, int32_t* gate, void* stream) {
  g_err[0] = 0;
  if (n <= 0 || p <= 0) return fail(MW_EINVAL, "%sbad sizes n=%lld p=%lld", "", n, p);
  if (!X0 || !C || !T || !Y || !scratch || !stats || !gate)
    return fail(MW_EINVAL, "%sNULL argument");
  if (ldx < p || ldc < p || ldt < p || ldy < p) return fail(MW_EINVAL, "%sbad leading dimension");
  if (!(tol >= 0.0)) return fail(MW_EINVAL, "%stolerance must be >= 0 (not NaN)");
  return hip_status(mw_launch_readout_check(n, p, X0, ldx, C, ldc, T, ldt, Y, ldy, tol, scratch,
                                            stats, gate, (hipStream_t)stream),
                    "mw_zreadout_check");
}

struct mw_sweep_plan {
  int nprob = 0, n_mesh = 0, gemm_variant = 0, max_ports = 0, max_con = 0, max_nse = 0;
  int64_t max_scatter = 0, n_items = 0, n_se_items = 0, n_groups = 0;
  int64_t n_plain = 0;             // fill items before those of extended-kernel meshes
  double pivot_bound = 32.0;
  mw_sweep_prob* dP = nullptr;
  mw_ports* dports = nullptr;
  mw_mesh* dmeshes = nullptr;
  mw_kgroup* d_groups = nullptr;   // frequency groups of the fill
  int32_t* d_group_prob = nullptr; // problems of each group, in wavenumber order
  int32_t* d_item_group = nullptr; // (group, tile pair) work items of the fill
  int32_t* d_item_pair = nullptr;
  int32_t* d_se_items = nullptr;   // (problem, 256-thread block) of the same-edge tables
  uint32_t* d_work = nullptr;      // per-step work lists of the factorisation
  int32_t* d_next = nullptr;       // claim counter of the persistent form
  std::vector<int64_t> work_off;
  std::vector<mw_sweep_prob> hp;
};

static void plan_free(mw_sweep_plan* p) {
  if (!p) return;
  (void)hipFree(p->dP);
  (void)hipFree(p->dports);
  (void)hipFree(p->dmeshes);
  (void)hipFree(p->d_groups);
  (void)hipFree(p->d_group_prob);
  (void)hipFree(p->d_item_group);
  (void)hipFree(p->d_item_pair);
  (void)hipFree(p->d_se_items);
  (void)hipFree(p->d_work);
  (void)hipFree(p->d_next);
  delete p;
}

// Frequency groups of the batched fill: the problems of one mesh (with one
// tile list and symmetry), in ascending wavenumber, cut into runs of at most
// kMaxRun whose wavenumbers step evenly to within 4 ulp.  The fill advances
// exp(-j k R) from one wavenumber of a run to the next by a fixed rotation,
// so a run's length also bounds the rounding that the rotations accumulate.
static constexpr int kMaxRun = 16;

static void sweep_kgroups(const std::vector<mw_sweep_prob>& hp, std::vector<mw_kgroup>& groups,
                          std::vector<int32_t>& gprob) {
  std::vector<int32_t> order(hp.size());
  for (size_t i = 0; i < hp.size(); ++i) order[i] = (int32_t)i;
  auto key_less = [&](int32_t a, int32_t b) {
    const mw_sweep_prob &x = hp[a], &y = hp[b];
    if (x.mesh != y.mesh) return x.mesh < y.mesh;
    if (x.tile_pairs != y.tile_pairs) return x.tile_pairs < y.tile_pairs;
    if (x.symmetric != y.symmetric) return x.symmetric < y.symmetric;
    if (x.n_pairs != y.n_pairs) return x.n_pairs < y.n_pairs;
    return x.k < y.k;
  };
  std::stable_sort(order.begin(), order.end(), key_less);
  auto same_mesh = [&](int32_t a, int32_t b) {
    const mw_sweep_prob &x = hp[a], &y = hp[b];
    return x.mesh == y.mesh && x.tile_pairs == y.tile_pairs && x.symmetric == y.symmetric &&
           x.n_pairs == y.n_pairs;
  };
  const double tol = 4.0 * 2.220446049250313e-16;
  size_t i = 0;
  while (i < order.size()) {
    mw_kgroup g;
    g.mesh = hp[order[i]].mesh;
    g.first = (int32_t)gprob.size();
    g.pad = 0;
    g.k0 = hp[order[i]].k;
    g.dk = 0.0;
    size_t j = i + 1;
    if (j < order.size() && same_mesh(order[i], order[j])) {
      g.dk = hp[order[j]].k - g.k0;
      ++j;
      while (j < order.size() && (int)(j - i) < kMaxRun && same_mesh(order[i], order[j])) {
        const double kj = hp[order[j]].k;
        if (std::fabs(kj - (g.k0 + (double)(j - i) * g.dk)) > tol * std::fabs(kj)) break;
        ++j;
      }
    }
    g.nk = (int32_t)(j - i);
    for (size_t t = i; t < j; ++t) gprob.push_back(order[t]);
    groups.push_back(g);
    i = j;
  }
}

static hipError_t sweep_fill(const mw_sweep_plan* p, double eps, double mu, hipStream_t st) {
  return mw_launch_batch_fill(p->dP, p->dmeshes, p->nprob, p->d_se_items, p->n_se_items,
                              p->d_groups, p->d_group_prob, p->d_item_group, p->d_item_pair,
                              p->n_items, p->n_plain, eps, mu, st);
}

int mw_sweep_plan_create(int32_t n_mesh, const mw_mesh* meshes, const mw_ports* ports,
                         int32_t nprob, const mw_sweep_prob* probs, int32_t gemm_variant,
                         double pivot_bound, mw_sweep_plan** out) {
  g_err[0] = 0;
  if (!out) return fail(MW_EINVAL, "%splan output is NULL");
  *out = nullptr;
  if (!(pivot_bound == pivot_bound)) return fail(MW_EINVAL, "%spivot threshold is NaN");
  if (n_mesh <= 0 || nprob <= 0 || !meshes || !ports || !probs)
    return fail(MW_EINVAL, "%sempty sweep");
  if (nprob > 65536) return fail(MW_ERANGE, "%smore than 65536 problems in one sweep");
  if (gemm_variant < 0 || gemm_variant > 5)
    return fail(MW_EINVAL, "%sgemm_variant %lld: expected 0..5", "", (long long)gemm_variant);
  for (int i = 0; i < n_mesh; ++i) {
    int rc = check_mesh(&meshes[i]);
    if (rc) return rc;
    if (meshes[i].n_qp != 4)
      return fail(MW_EINVAL, "%smesh %lld: the batched fill is built for n_qp_pair = 4", "", i);
    if (ports[i].n_ports < 1 || ports[i].n_ports > 32 || ports[i].n_con < 0 ||
        ports[i].n_con > 32 || ports[i].n_ports + ports[i].n_con > 64)
      return fail(MW_ERANGE, "%smesh %lld: ports/constraints beyond the batched limits", "", i);
  }
  mw_sweep_plan* p = new mw_sweep_plan();
  p->nprob = nprob;
  p->n_mesh = n_mesh;
  p->gemm_variant = gemm_variant;
  p->pivot_bound = pivot_bound;
  p->hp.assign(probs, probs + nprob);
  for (const auto& q : p->hp) {
    if (q.mesh < 0 || q.mesh >= n_mesh || q.n != meshes[q.mesh].n_basis ||
        q.ncols != ports[q.mesh].n_ports + ports[q.mesh].n_con || q.lda < q.n + q.ncols ||
        q.n > 4096 || !q.A || !q.ipiv || !q.work || !q.info || !q.Y || !q.se_tab || !q.tile_pairs) {
      plan_free(p);
      return fail(MW_EINVAL, "%sinconsistent sweep problem (mesh %lld, n %lld)", "", q.mesh, q.n);
    }
    const mw_ports& pt = ports[q.mesh];
    p->max_ports = std::max(p->max_ports, pt.n_ports);
    p->max_con = std::max(p->max_con, pt.n_con);
    p->max_scatter = std::max<int64_t>(p->max_scatter, (int64_t)q.ncols * pt.col_max);
    p->max_nse = std::max(p->max_nse, meshes[q.mesh].n_se);
  }
  std::stable_sort(p->hp.begin(), p->hp.end(),
                   [](const mw_sweep_prob& a, const mw_sweep_prob& b) { return a.n > b.n; });
  std::vector<uint32_t> work;
  if (gemm_variant < 2) mw_sweep_work(p->hp.data(), nprob, gemm_variant, work, p->work_off);
  std::vector<mw_kgroup> groups;
  std::vector<int32_t> gprob, ig, iq;
  sweep_kgroups(p->hp, groups, gprob);
  for (int ek = 0; ek < 2; ++ek) {
    if (ek == 1) p->n_plain = (int64_t)ig.size();
    for (size_t g = 0; g < groups.size(); ++g) {
      if ((meshes[groups[g].mesh].ek != 0) != (ek == 1)) continue;
      const int32_t np = p->hp[gprob[groups[g].first]].n_pairs;
      for (int32_t j = 0; j < np; ++j) {
        ig.push_back((int32_t)g);
        iq.push_back(j);
      }
    }
  }
  p->n_groups = (int64_t)groups.size();
  p->n_items = (int64_t)ig.size();
  std::vector<int32_t> se_items;
  for (int i = 0; i < nprob; ++i) {
    const int64_t nb = mw_same_edge_blocks(meshes[p->hp[i].mesh].n_se);
    for (int64_t b = 0; b < nb; ++b) {
      se_items.push_back(i);
      se_items.push_back((int32_t)b);
    }
  }
  p->n_se_items = (int64_t)se_items.size() / 2;
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&p->dP, sizeof(mw_sweep_prob) * nprob)) != hipSuccess ||
      (e = hipMalloc(&p->dports, sizeof(mw_ports) * n_mesh)) != hipSuccess ||
      (e = hipMalloc(&p->dmeshes, sizeof(mw_mesh) * n_mesh)) != hipSuccess ||
      (e = hipMalloc(&p->d_groups, sizeof(mw_kgroup) * std::max<size_t>(groups.size(), 1))) != hipSuccess ||
      (e = hipMalloc(&p->d_group_prob, sizeof(int32_t) * std::max<size_t>(gprob.size(), 1))) != hipSuccess ||
      (e = hipMalloc(&p->d_item_group, sizeof(int32_t) * std::max<int64_t>(p->n_items, 1))) != hipSuccess ||
      (e = hipMalloc(&p->d_item_pair, sizeof(int32_t) * std::max<int64_t>(p->n_items, 1))) != hipSuccess ||
      (e = hipMemcpy(p->dP, p->hp.data(), sizeof(mw_sweep_prob) * nprob, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->dports, ports, sizeof(mw_ports) * n_mesh, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->dmeshes, meshes, sizeof(mw_mesh) * n_mesh, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_groups, groups.data(), sizeof(mw_kgroup) * groups.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_group_prob, gprob.data(), sizeof(int32_t) * gprob.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (!ig.empty() &&
       (e = hipMemcpy(p->d_item_group, ig.data(), sizeof(int32_t) * ig.size(), hipMemcpyHostToDevice)) != hipSuccess) ||
      (!iq.empty() &&
       (e = hipMemcpy(p->d_item_pair, iq.data(), sizeof(int32_t) * iq.size(), hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMalloc(&p->d_se_items, sizeof(int32_t) * std::max<size_t>(se_items.size(), 2))) != hipSuccess ||
      (!se_items.empty() &&
       (e = hipMemcpy(p->d_se_items, se_items.data(), sizeof(int32_t) * se_items.size(), hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMalloc(&p->d_work, sizeof(uint32_t) * std::max<size_t>(work.size(), 1))) != hipSuccess ||
      (e = hipMalloc(&p->d_next, sizeof(int32_t) * 2)) != hipSuccess ||
      (!work.empty() &&
       (e = hipMemcpy(p->d_work, work.data(), sizeof(uint32_t) * work.size(), hipMemcpyHostToDevice)) != hipSuccess)) {
    plan_free(p);
    return hip_status(e, "mw_sweep_plan_create");
  }
  *out = p;
  return MW_OK;
}

int mw_sweep_run(mw_sweep_plan* p, double eps, double mu, const volatile int32_t* cancel,
                 void* stream) {
  g_err[0] = 0;
  if (!p) return fail(MW_EINVAL, "%splan is NULL");
  hipStream_t st = (hipStream_t)stream;
  if (cancel && *cancel) return fail(MW_EABORT, "%ssweep cancelled");
  hipError_t herr = sweep_fill(p, eps, mu, st);
  if (herr != hipSuccess) return hip_status(herr, "sweep fill");
  int rc;
  if (p->gemm_variant >= 2) {
    // one launch: the kernel reads the flag itself when it can (mapped or
    // device memory); an ordinary host flag is only read here, before it
    const int v = p->gemm_variant - 2;
    rc = mw_sweep_wg_run(p->dP, p->dports, p->nprob, v & 1, (v & 2) ? 1 : 2, p->pivot_bound,
                         p->d_next, device_view(cancel), st, &herr);
  } else {
    rc = mw_sweep_lu_run(p->dP, p->dports, p->hp.data(), p->nprob, p->max_ports, p->max_con,
                         p->max_scatter, p->gemm_variant, p->pivot_bound, p->d_work,
                         p->work_off.data(), cancel, st, &herr);
  }
  if (rc == MW_EHIP) return hip_status(herr, "sweep solve");
  if (rc == MW_EABORT) return fail(MW_EABORT, "%ssweep cancelled");
  return rc;
}

int mw_sweep_fill(mw_sweep_plan* p, double eps, double mu, void* stream) {
  g_err[0] = 0;
  if (!p) return fail(MW_EINVAL, "%splan is NULL");
  return hip_status(sweep_fill(p, eps, mu, (hipStream_t)stream), "sweep fill");
}

// --- block-low-rank operator --------------------------------------------
static_assert(sizeof(mw_hblock) == 72, "mw_hblock layout (the Python binding mirrors it)");
static_assert(sizeof(mw_hm_titem) == 24 && sizeof(mw_hm_ritem) == 24, "work item layouts");

static int check_entry_args(const mw_mesh* mesh, const int32_t* wing_seg,
                            const int32_t* wing_slot, double omega, double eps, double mu) {
  int rc = check_mesh(mesh);
  if (rc) return rc;
  if (!wing_seg || !wing_slot) return fail(MW_EINVAL, "%swing tables are NULL");
  if (!(omega > 0.0) || !(eps > 0.0) || !(mu > 0.0))
    return fail(MW_EINVAL, "%somega, eps and mu must be positive");
  return MW_OK;
}

int mw_zentries(const mw_mesh* mesh, const int32_t* wing_seg, const int32_t* wing_slot, double k,
                double omega, double eps, double mu, const int32_t* rows, const int32_t* cols,
                int64_t n, int32_t same_edge_table, void* out, void* stream) {
  g_err[0] = 0;
  int rc = check_entry_args(mesh, wing_seg, wing_slot, omega, eps, mu);
  if (rc) return rc;
  if (n < 0) return fail(MW_EINVAL, "%snegative entry count");
  if (n > 0 && (!rows || !cols || !out)) return fail(MW_EINVAL, "%sNULL argument");
  return hip_status(mw_launch_zentries(*mesh, wing_seg, wing_slot, k, omega, eps, mu, rows, cols, n,
                                       same_edge_table, out, (hipStream_t)stream),
                    "mw_zentries");
}

int mw_hm_fill_dense(const mw_mesh* mesh, const int32_t* wing_seg, const int32_t* wing_slot,
                     double k, double omega, double eps, double mu, const int32_t* rows_pool,
                     const int32_t* cols_pool, const mw_hblock* blocks, const int64_t* items,
                     int64_t n_items, int32_t same_edge_table, void* store, void* stream) {
  g_err[0] = 0;
  int rc = check_entry_args(mesh, wing_seg, wing_slot, omega, eps, mu);
  if (rc) return rc;
  if (n_items < 0 || n_items > 0x7fffffffLL) return fail(MW_ERANGE, "%sbad work item count");
  if (n_items > 0 && (!rows_pool || !cols_pool || !blocks || !items || !store))
    return fail(MW_EINVAL, "%sNULL argument");
  return hip_status(mw_launch_hm_dense(*mesh, wing_seg, wing_slot, k, omega, eps, mu, rows_pool,
                                       cols_pool, blocks, items, n_items, same_edge_table, store,
                                       (hipStream_t)stream),
                    "mw_hm_fill_dense");
}

int mw_hm_aca(const mw_mesh* mesh, const int32_t* wing_seg, const int32_t* wing_slot, double k,
              double omega, double eps, double mu, const int32_t* rows_pool,
              const int32_t* cols_pool, mw_hblock* blocks, const int32_t* list, int32_t n_list,
              int32_t n_large, double tol, void* store, void* work, void* stream) {
  g_err[0] = 0;
  int rc = check_entry_args(mesh, wing_seg, wing_slot, omega, eps, mu);
  if (rc) return rc;
  if (n_list < 0 || n_large < 0 || n_large > n_list)
    return fail(MW_EINVAL, "%sbad block counts");
  if (!(tol > 0.0)) return fail(MW_EINVAL, "%sACA tolerance must be positive");
  if (n_list > 0 && (!rows_pool || !cols_pool || !blocks || !list || !store || !work))
    return fail(MW_EINVAL, "%sNULL argument");
  return hip_status(mw_launch_hm_aca(*mesh, wing_seg, wing_slot, k, omega, eps, mu, rows_pool,
                                     cols_pool, blocks, list, n_list, n_large, tol, store, work,
                                     (hipStream_t)stream),
                    "mw_hm_aca");
}

int mw_hm_apply(const mw_hblock* blocks, const void* store, const mw_hm_titem* t_items,
                int32_t n_t, const int32_t* t_rows, const int32_t* colpos,
                const mw_hm_ritem* r_items, int32_t n_r, const int32_t* r_ent,
                const int32_t* sigma, const void* X, int64_t ldx, int32_t nrhs, void* T,
                void* Y, int64_t ldy, void* stream) {
  g_err[0] = 0;
  if (nrhs < 0 || n_t < 0 || n_r < 0) return fail(MW_EINVAL, "%snegative size");
  if (nrhs > 0) {
    if (!blocks || !store || !X || !Y || (n_r > 0 && (!r_items || !r_ent)) ||
        (n_t > 0 && (!t_items || !t_rows || !colpos || !T)))
      return fail(MW_EINVAL, "%sNULL argument");
    if (ldx < nrhs || ldy < nrhs) return fail(MW_EINVAL, "%sbad leading dimension");
  }
  return hip_status(mw_launch_hm_apply(blocks, store, t_items, n_t, t_rows, colpos, r_items, n_r,
                                       r_ent, sigma, X, ldx, nrhs, T, Y, ldy, (hipStream_t)stream),
                    "mw_hm_apply");
}

int mw_hm_invert(const mw_hblock* blocks, const int32_t* list, int32_t n_list, void* store,
                 void* work, int32_t* info, void* stream) {
  g_err[0] = 0;
  if (n_list < 0) return fail(MW_EINVAL, "%snegative block count");
  if (n_list > 0 && (!blocks || !list || !store || !work || !info))
    return fail(MW_EINVAL, "%sNULL argument");
  return hip_status(mw_launch_hm_invert(blocks, list, n_list, store, work, info,
                                        (hipStream_t)stream),
                    "mw_hm_invert");
}

int mw_sweep_plan_destroy(mw_sweep_plan* p) {
  plan_free(p);
  return MW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Sommerfeld ground
// ---------------------------------------------------------------------------
namespace {
int check_grid(const mw_somm_grid* G) {
  if (!G) return fail(MW_EINVAL, "%sgrid is NULL");
  if (G->n_reg != 4 && G->n_reg != 6)
    return fail(MW_EINVAL, "%sgrid must have 4 or 6 regions, got %lld", "", G->n_reg);
  if (!G->vals) return fail(MW_EINVAL, "%sgrid values are NULL");
  if (!(G->r1_max > 0.0) || !(G->r_break > 0.0) || !(G->r_near > 0.0))
    return fail(MW_EINVAL, "%sgrid radii must be positive");
  for (int g = 0; g < G->n_reg; ++g) {
    if (G->n_r[g] < 4 || G->n_th[g] < 4)
      return fail(MW_EINVAL, "%sregion %lld has fewer than 4 x 4 nodes", "", g);
    if (!(G->dr[g] > 0.0) || !(G->dth[g] > 0.0))
      return fail(MW_EINVAL, "%sregion %lld spacing must be positive", "", g);
    if (G->off[g] < 0) return fail(MW_EINVAL, "%sregion %lld offset is negative", "", g);
  }
  return MW_OK;
}
}  // namespace

int mw_somm_six_integrals(double eps_re, double eps_im, double k2, const double* rho,
                          const double* h, int64_t n, double rtol, int32_t form, void* out,
                          const volatile int32_t* cancel, void* stream) {
  g_err[0] = 0;
  if (n < 0) return fail(MW_EINVAL, "%snode count is negative");
  if (n > 0 && (!rho || !h || !out)) return fail(MW_EINVAL, "%snode or output array is NULL");
  if (!(k2 > 0.0)) return fail(MW_EINVAL, "%sk2 must be positive");
  if (!(rtol > 0.0)) return fail(MW_EINVAL, "%srtol must be positive");
  if (form < 0 || form > 2) return fail(MW_EINVAL, "%sform must be 0 (auto), 1 (J) or 2 (H)");
  if (!(eps_re == eps_re) || !(eps_im == eps_im)) return fail(MW_EINVAL, "%seps is NaN");
  const volatile int32_t* dcancel = nullptr;
  int rc = device_cancel(cancel, &dcancel);
  if (rc) return rc;
  return hip_status(mw_launch_somm_six(eps_re, eps_im, k2, rho, h, n, rtol, form,
                                       (double2*)out, dcancel, (hipStream_t)stream),
                    "mw_somm_six_integrals");
}

int mw_somm_surfaces(double eps_re, double eps_im, double k2, double omega, double mu,
                     const double* R1, const double* theta, int64_t n, double rtol,
                     void* out, const volatile int32_t* cancel, void* stream) {
  g_err[0] = 0;
  if (n < 0) return fail(MW_EINVAL, "%snode count is negative");
  if (n > 0 && (!R1 || !theta || !out)) return fail(MW_EINVAL, "%snode or output array is NULL");
  if (!(k2 > 0.0) || !(omega > 0.0) || !(mu > 0.0))
    return fail(MW_EINVAL, "%sk2, omega and mu must be positive");
  if (!(rtol > 0.0)) return fail(MW_EINVAL, "%srtol must be positive");
  if (!(eps_re == eps_re) || !(eps_im == eps_im)) return fail(MW_EINVAL, "%seps is NaN");
  const volatile int32_t* dcancel = nullptr;
  int rc = device_cancel(cancel, &dcancel);
  if (rc) return rc;
  return hip_status(mw_launch_somm_surfaces(eps_re, eps_im, k2, omega, mu, R1, theta, n, rtol,
                                            (double2*)out, dcancel, (hipStream_t)stream),
                    "mw_somm_surfaces");
}

int mw_somm_eval(const mw_somm_grid* grid, const double* R1, const double* theta, int64_t n,
                 void* out, void* stream) {
  g_err[0] = 0;
  int rc = check_grid(grid);
  if (rc) return rc;
  if (n < 0) return fail(MW_EINVAL, "%spoint count is negative");
  if (n > 0 && (!R1 || !theta || !out)) return fail(MW_EINVAL, "%spoint or output array is NULL");
  return hip_status(mw_launch_somm_eval(*grid, R1, theta, n, (double2*)out, (hipStream_t)stream),
                    "mw_somm_eval");
}

int mw_ground_fields(const mw_somm_grid* grid, double ground_z, double k, int32_t n_pts,
                     const double* pts, int32_t n_src, const double* mid, const void* moment,
                     int64_t ldm, int32_t p, int32_t accumulate, void* out, int64_t ldo,
                     void* stream) {
  g_err[0] = 0;
  int rc = check_grid(grid);
  if (rc) return rc;
  if (n_pts < 0 || n_src < 0 || p < 1) return fail(MW_EINVAL, "%sbad point, source or column count");
  if (!(k > 0.0)) return fail(MW_EINVAL, "%sk must be positive");
  if (ldm < 3LL * p || ldo < 3LL * p)
    return fail(MW_EINVAL, "%sldm %lld / ldo %lld below 3 p", "", ldm, ldo);
  if ((n_pts > 0 && (!pts || !out)) || (n_src > 0 && (!mid || !moment)))
    return fail(MW_EINVAL, "%sNULL array");
  return hip_status(mw_launch_ground_fields(*grid, ground_z, k, n_pts, pts, n_src, mid,
                                            (const double2*)moment, ldm, p, accumulate,
                                            (double2*)out, ldo, (hipStream_t)stream),
                    "mw_ground_fields");
}

int mw_fill_ground(const mw_mesh* mesh, double k, double omega, double eps, double mu,
                   double ground_z, double c2_re, double c2_im, const mw_somm_grid* grid,
                   int32_t n_qp_sommerfeld, void* Z, int64_t ldz, const int32_t* tile_pairs,
                   int32_t n_pairs, int32_t symmetric, void* stream) {
  g_err[0] = 0;
  int rc = check_mesh_ground(mesh);
  if (rc) return rc;
  if (grid && (rc = check_grid(grid))) return rc;
  if (n_qp_sommerfeld < 1 || n_qp_sommerfeld > 8)
    return fail(MW_EINVAL, "%sn_qp_sommerfeld %lld not built (1 .. 8)", "", n_qp_sommerfeld);
  if (!Z) return fail(MW_EINVAL, "%sZ is NULL");
  if (ldz < mesh->n_basis) return fail(MW_EINVAL, "%sldz %lld < n_basis %lld", "", ldz, mesh->n_basis);
  if (!(k > 0.0) || !(omega > 0.0)) return fail(MW_EINVAL, "%sk and omega must be positive");
  if (symmetric && (!tile_pairs || n_pairs <= 0))
    return fail(MW_EINVAL, "%ssymmetric fill needs the tile-pair list");
  return hip_status(mw_launch_fill_ground(*mesh, k, omega, eps, mu, ground_z, c2_re, c2_im, grid,
                                          nullptr, n_qp_sommerfeld, (double2*)Z, ldz, tile_pairs,
                                          n_pairs,
                                          symmetric != 0, (hipStream_t)stream),
                    "mw_fill_ground");
}

int mw_fill_ground_refl(const mw_mesh* mesh, double k, double omega, double eps, double mu,
                        double ground_z, double eps_t_re, double eps_t_im, double phi_c0_re,
                        double phi_c0_im, double phi_c1_re, double phi_c1_im, void* Z,
                        int64_t ldz, const int32_t* tile_pairs, int32_t n_pairs,
                        int32_t symmetric, void* stream) {
  g_err[0] = 0;
  int rc = check_mesh_ground(mesh);
  if (rc) return rc;
  if (!Z) return fail(MW_EINVAL, "%sZ is NULL");
  if (ldz < mesh->n_basis) return fail(MW_EINVAL, "%sldz %lld < n_basis %lld", "", ldz, mesh->n_basis);
  if (!(k > 0.0) || !(omega > 0.0)) return fail(MW_EINVAL, "%sk and omega must be positive");
  if (!std::isfinite(eps_t_re) || !std::isfinite(eps_t_im))
    return fail(MW_EINVAL, "%sground permittivity is not finite");
  if (symmetric && (!tile_pairs || n_pairs <= 0))
    return fail(MW_EINVAL, "%ssymmetric fill needs the tile-pair list");
  const double refl[6] = {eps_t_re, eps_t_im, phi_c0_re, phi_c0_im, phi_c1_re, phi_c1_im};
  return hip_status(mw_launch_fill_ground(*mesh, k, omega, eps, mu, ground_z, 1.0, 0.0, nullptr,
                                          refl, 3, (double2*)Z, ldz, tile_pairs, n_pairs,
                                          symmetric != 0, (hipStream_t)stream),
                    "mw_fill_ground_refl");
}

int mw_somm_surfaces_varied(const double* eps, double k2, double omega, double mu,
                            const double* R1, const double* theta, int64_t n, double rtol,
                            void* out, const volatile int32_t* cancel, void* stream) {
  g_err[0] = 0;
  if (n < 0) return fail(MW_EINVAL, "%snode count is negative");
  if (n > 0 && (!eps || !R1 || !theta || !out))
    return fail(MW_EINVAL, "%snode, permittivity or output array is NULL");
  if (!(k2 > 0.0) || !(omega > 0.0) || !(mu > 0.0))
    return fail(MW_EINVAL, "%sk2, omega and mu must be positive");
  if (!(rtol > 0.0)) return fail(MW_EINVAL, "%srtol must be positive");
  const volatile int32_t* dcancel = nullptr;
  int rc = device_cancel(cancel, &dcancel);
  if (rc) return rc;
  return hip_status(mw_launch_somm_surfaces_varied(eps, k2, omega, mu, R1, theta, n, rtol,
                                                   (double2*)out, dcancel, (hipStream_t)stream),
                    "mw_somm_surfaces_varied");
}

// ---------------------------------------------------------------------------
// learned surrogate
// ---------------------------------------------------------------------------
namespace {
int check_mlp(const mw_mlp* net, bool train) {
  if (!net || !net->params || !net->w16 || !net->wt16)
    return fail(MW_EINVAL, "%snetwork buffers are NULL");
  if (train && (!net->grads || !net->adam_m || !net->adam_v))
    return fail(MW_EINVAL, "%straining buffers are NULL");
  return MW_OK;
}
}  // namespace

int64_t mw_mlp_param_count(void) { return mw_mlp_n_params(); }
int64_t mw_mlp_w16_count(void) { return mw_mlp_n_w16(); }
int64_t mw_mlp_workspace_bytes(int64_t n) { return n > 0 ? mw_mlp_ws_bytes((n + 63) / 64 * 64) : 0; }

int mw_mlp_sync_weights(const mw_mlp* net, void* stream) {
  g_err[0] = 0;
  int rc = check_mlp(net, false);
  if (rc) return rc;
  return hip_status(mw_mlp_launch_sync(*net, (hipStream_t)stream), "mw_mlp_sync_weights");
}

int mw_mlp_forward(const mw_mlp* net, const float* x, int64_t n, float* out, void* ws,
                   void* stream) {
  g_err[0] = 0;
  int rc = check_mlp(net, false);
  if (rc) return rc;
  if (n <= 0) return fail(MW_EINVAL, "%sbatch must be positive");
  if (!x || !out || !ws) return fail(MW_EINVAL, "%sinput, output or workspace is NULL");
  return hip_status(mw_mlp_launch_forward(*net, x, n, out, ws, (hipStream_t)stream),
                    "mw_mlp_forward");
}

int mw_mlp_backward(const mw_mlp* net, const float* x, const float* target, int64_t n,
                    float* loss, void* ws, void* stream) {
  g_err[0] = 0;
  int rc = check_mlp(net, true);
  if (rc) return rc;
  if (n <= 0) return fail(MW_EINVAL, "%sbatch must be positive");
  if (!x || !target || !ws) return fail(MW_EINVAL, "%sinput, target or workspace is NULL");
  return hip_status(mw_mlp_launch_backward(*net, x, target, n, loss, ws, (hipStream_t)stream),
                    "mw_mlp_backward");
}

int mw_mlp_adam_step(const mw_mlp* net, float lr, float beta1, float beta2, float eps,
                     int32_t step, void* stream) {
  g_err[0] = 0;
  int rc = check_mlp(net, true);
  if (rc) return rc;
  if (step < 1) return fail(MW_EINVAL, "%sAdam step counts from 1");
  if (!(lr >= 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return fail(MW_EINVAL, "%sAdam hyper-parameters out of range");
  return hip_status(mw_mlp_launch_adam(*net, lr, beta1, beta2, eps, step, (hipStream_t)stream),
                    "mw_mlp_adam_step");
}

int mw_mlp_grid_values(const mw_mlp* net, double eps_re, double eps_im, const double* R1,
                       const double* theta, const int32_t* idx, int64_t n, void* vals,
                       double value_re, double value_im, void* ws, void* stream) {
  g_err[0] = 0;
  int rc = check_mlp(net, false);
  if (rc) return rc;
  if (n < 0) return fail(MW_EINVAL, "%snode count is negative");
  if (n > 0 && (!R1 || !theta || !idx || !vals || !ws))
    return fail(MW_EINVAL, "%snode, index, value or workspace array is NULL");
  return hip_status(mw_mlp_launch_grid(*net, eps_re, eps_im, R1, theta, idx, n, (double2*)vals,
                                       value_re, value_im, ws, (hipStream_t)stream),
                    "mw_mlp_grid_values");
}
