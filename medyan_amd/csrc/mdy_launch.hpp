// mdy_launch.hpp -- host-side launchers shared between translation units.
#pragma once
#include "mdy_internal.hpp"

namespace mdy {

// mdy_eval.hip
int launch_energy(Ctx &c, const double *d_x, double *d_out, hipStream_t st);
int launch_forces(Ctx &c, const double *d_x, double *d_force, int accumulate, hipStream_t st);
// allocation-free energy: scratch in a file-scope device array of mdy_eval.hip
int resolve_noalloc_scratch(int device);
size_t noalloc_scratch_bytes();
int launch_energy_noalloc(Ctx &c, const double *d_x, double *d_out, hipStream_t st);
// mdy_nl.hip
int build_neighbors(Ctx &c, const double *d_x);
int exclusive_scan(Ctx &c, const int *in, int *out, int n, hipStream_t st);  // out has n+1 entries
// mdy_cg.hip
int run_minimize(Ctx &c, const mdy_cg_params *p, double *d_x, double *d_force_out, mdy_min_result *res);
// CGMethod::searchDir after the context's last minimization, to host memory (3 n_beads doubles)
int final_search_dir(Ctx &c, double *h_out);
int run_minimize_batch(Ctx **cs, int n, const mdy_cg_params *p, double *const *d_x, double *const *d_force_out,
                       mdy_min_result *res, double *net_ms, int threads_per_block, const double *weights,
                       int first_pass_calls, hipStream_t st, double *total_ms);
// mdy_hvp.hip
// sync = false: enqueue only; a device error stays in the context's error
// word for sync_errors (or the next call that waits)
int launch_hvp(Ctx &c, const double *d_x, const double *d_v, double *d_hv, hipStream_t st, bool sync = true);
int sync_errors(Ctx &c, hipStream_t st);
int launch_energy_jvp(Ctx &c, const double *d_x, const double *d_w, double *d_out, hipStream_t st);
// mdy_quad.hip: quadrature mode (MDY_COPLANAR_QUADRATURE)
bool quad_on(const Ctx &c);
int quad_blocks(const Ctx &c);   // block records the quadrature energy / JVP kernels append
// flags the pairs evaluated by quadrature at d_x and points D.evq at the flags
int quad_classify(Ctx &c, Dev &D, const double *d_x, hipStream_t st);
// what: 0 energy (block records from rec0), 1 forces (slots), 2 H v along d_v (slots),
// 3 grad E_EV . d_v (block records from rec0)
int launch_quad(Ctx &c, const Dev &D, int what, const double *d_x, const double *d_v, double *out, int rec0,
                hipStream_t st);
int quad_pair_count(Ctx &c, const double *d_x, hipStream_t st, int *count);
// mdy_aux.hip
int launch_stretch_forces(Ctx &c, const double *d_x, double *d_linker, double *d_motor, double *d_branch,
                          hipStream_t st);
int launch_load_forces(Ctx &c, const double *d_x, int ntypes, const double *d_mon_size, const int *d_n_mon,
                       int max_mon, double *d_lp, double *d_lm, hipStream_t st);

template <class T>
int ensure(DBuf<T> &b, size_t n) {
    if (n == 0) n = 1;
    if (b.n >= n) return MDY_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    size_t cap = n + n / 4 + 16;
    if (hipMalloc((void **)&b.p, cap * sizeof(T)) != hipSuccess) return MDY_ERR_HIP;
    b.n = cap;
    return MDY_OK;
}
template <class T>
void release(DBuf<T> &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
}

#define MDY_TRY(expr)                    \
    do {                                 \
        int _rc = (expr);                \
        if (_rc != MDY_OK) return _rc;   \
    } while (0)
#define MDY_HIP(expr)                                        \
    do {                                                     \
        if ((expr) != hipSuccess) return MDY_ERR_HIP;        \
    } while (0)

}  // namespace mdy
