// Host-side state of one 3D frame transient (frame3d.hip) and its modal
// analysis (modal.hip); not part of the C ABI.
#pragma once

#include <cstdint>
#include <vector>

struct sgpu_ctx;

namespace sgpu {

constexpr int kF3MaxRet = 12;       // free retained DOFs one element reaches
constexpr int kF3ShellMaxRet = 24;  // free retained DOFs one shell reaches
constexpr int kF3SolveThreads = 1024;

struct Frame3dHost {
    int n_nodes = 0, total = 0, n = 0, hbw = 0, W = 0, n_el = 0;
    double dt = 0, gamma = 0, beta = 0, tol = 0, alpha_m = 0, beta_k = 0, beta_k_init = 0;
    int max_iters = 0, n_patterns = 0;
    // records: ground-motion instances of the frame, each with its own
    // elements (section states) and series; per-record device arrays hold
    // R slices
    int R = 0;
    std::vector<int32_t> rec_elem_ids;                 // R x n_el
    std::vector<std::vector<double>> rec_ts;           // R x n_patterns series
    std::vector<double> rec_ts_dt, rec_ts_factor;      // R x n_patterns
    std::vector<int32_t> rec_steps, rec_status;        // completed steps; 0 running, > 0 stopped, < 0 failed
    std::vector<uint8_t> elem_used;                    // per sgpu_elem3d element: claimed by a record
    std::vector<int32_t> list_on_device;               // the record list the kernels read now
    bool allocated = false, initialized = false;
    int32_t *d_list = nullptr, *d_list_ids = nullptr, *d_row_of = nullptr;
    double *d_ag = nullptr;  // R x 3 ground accelerations of the step
    // device tables
    int32_t *d_free_idx = nullptr, *d_row_offsets = nullptr, *d_dof_pool = nullptr, *d_slave = nullptr;
    double *d_coeff_pool = nullptr;
    int32_t *d_elem_nodes = nullptr, *d_ret = nullptr, *d_ret_n = nullptr;
    double *d_T = nullptr;                                 // n_el x 12 x 12: T_e[a][a']
    int32_t *d_band_off = nullptr, *d_band_src = nullptr;  // per band slot: k_hat entries
    double *d_mass_band = nullptr, *d_kinit_band = nullptr;
    int32_t *d_dof_off = nullptr, *d_dof_src = nullptr;    // per free DOF: (e, a') entries
    int32_t *d_m_off = nullptr, *d_m_col = nullptr;        // reduced mass, CSR
    double *d_m_val = nullptr;
    double *d_load = nullptr;                              // n_patterns x n: T^T (-M iota_p)
    // elastic shells: stiffness formed and collapsed once; per record the
    // element-level products k_hat u and k_hat v of every iteration
    int n_sh = 0;
    int32_t *d_sh_ret = nullptr, *d_sh_ret_n = nullptr;    // n_sh x 24 free retained DOFs
    double *d_sh_khat = nullptr;                           // n_sh x 24 x 24 collapsed stiffness
    double *d_sh_f = nullptr, *d_sh_y = nullptr;           // R x n_sh x 24
    // device state
    double *d_u = nullptr;  // total
    double *d_uf, *d_vf, *d_af, *d_un, *d_vn, *d_an, *d_vt, *d_at, *d_R, *d_du;
    double *d_u12 = nullptr, *d_k = nullptr, *d_f = nullptr, *d_khat = nullptr, *d_kinit_hat = nullptr,
           *d_fhat = nullptr, *d_yk = nullptr, *d_yi = nullptr;
    int32_t *d_status = nullptr;
    double *d_band = nullptr;
    double *d_bwork = nullptr;  // R x band_piv_work_doubles when the register solve cannot hold the band
    int solver = 0;             // 0: the reference's row swaps (band LU / dense fallback); 1: no row swaps
    double *d_scalar = nullptr;  // [0] residual norm^2, [1] max element status, [2] solve status
    // recorder: chosen DOFs after every converged step, kept on the device
    int n_rec = 0, rec_cap = 0;
    int32_t *d_rec_dofs = nullptr;
    double *d_rec = nullptr;
    // host copies the modal analysis reads: free DOF list, expansion
    // tables, constraint flags and the reduced mass rows (CSR)
    std::vector<int32_t> h_free_list, h_constrained, h_ro, h_dp, h_slave, h_moff, h_mcol;
    std::vector<double> h_cp, h_mval;
};

// the initial assembly of every record (k_hat at rest kept as the initial
// tangent), once per frame
int frame3d_initialize(::sgpu_ctx *ctx, Frame3dHost &h);

}  // namespace sgpu
