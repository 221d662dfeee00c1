// Internal host-side types of libmxhip.so (not part of the C ABI).
#pragma once

#include <gmp.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mx.h"
#include "mx_dbl.h"
#include "mx_dev.h"

namespace mx {

struct Mpf {
    mpf_t v;
};

// Host state of one frame: grid, reference orbit and BLA table in the engine's
// formats (mandel_perturbation.h:193-206, 295-341).
struct Frame {
    Frame();
    ~Frame();
    Frame(const Frame&) = delete;
    Frame& operator=(const Frame&) = delete;

    int prepare(const mx_request& r, std::string& err);

    // host-mapped cancellation flag of the calling context (null: not cancellable); the
    // long host loops (reference orbit, nucleus search) poll it, as the engine's
    // createRef polls _flag_halt (mandel_perturbation.cpp:4140), and prepare() then
    // returns MX_E_CANCELLED
    const volatile int* halt = nullptr;
    bool halted() const { return halt && *halt; }
    void fill_view(mx_frame_view& v) const;

    // inputs
    int W = 0, H = 0, max_iter = 0;
    bool ede = false, use_bla = true, interior = true;
    float escape_radius = 1e8f;
    double esc2 = 0.0;
    mp_bitcnt_t prec = 64;
    int bla_minlevel = 3;
    int sub = 1;      // output cells per pixel side
    bool ss = false;  // adaptive supersampling (sub > 1 and MX_SUPER_SAMPLING)

    // grid
    Mpf c0_re, c0_im, dx, dy;
    Fe dxfe{0.0, 0}, dyfe{0.0, 0};

    // reference
    Mpf ref_re, ref_im, tail_re, tail_im;
    int ref_it = 0;
    int period = 0;        // > 0: the reference is a period-`period` nucleus, its orbit one period
    bool subpixel = false; // a nucleus in view spans < 4 pixels: full reference with BLA off
    bool ref_escaped = false, ref_virtual = false, cutoff_sensitive = false;
    double ref_x = 0.0, ref_y = 0.0;
    int ref_pixel_n = -1;
    double ref_pixel_value = -2.0;
    std::vector<double> zf, zfe_m;
    std::vector<int32_t> zfe_e;

    // BLA
    std::vector<int32_t> lev_off, lev_size;
    std::vector<double> bla_r2, bla_coef;
    double rmax2 = 0.0, dcmax = 0.0;

    // per-pixel dc
    std::vector<double> dcol_m, drow_m;
    std::vector<int32_t> dcol_e, drow_e;

    double ref_seconds = 0.0, bla_seconds = 0.0, prepare_seconds = 0.0;

    // shallow direct path (scale <= 1e6): c0, dx, dy truncated to double (mpf_get_ld)
    bool shallow = false;
    double sh_c0r = 0.0, sh_c0i = 0.0, sh_dx = 0.0, sh_dy = 0.0;

    // plain-double path (1e6 < scale <= 1e280): pixel references, the series
    // approximation, the glitch metric's orbit magnitudes and double dc tables
    bool dbl = false;
    bool escalated = false; // 1e150 < scale <= 1e280 frame moved to the floatexp path
    int ref_i = 0, ref_j = 0;
    int ref_u = 0, ref_v = 0;      // the reference's sub-sample offsets (row, column)
    std::vector<double> z_m3;
    MxC sa_A[MX_SA_N] = {};
    MxC sa_B[MX_SA_N] = {};        // series of the derivative delta (EDE)
    std::vector<double> dfr, dfi;  // reference derivative D (EDE)
    std::vector<double> bla_der;   // BLA derivative couplings C, E (EDE)
    MxC sa_delta = {0.0, 0.0};
    int sa_it = 0, sa_order = 0;
    bool sa_flag = false;
    std::vector<double> dcol_d, drow_d;
    int n_refs = 0;
    // Build the reference at pixel (pi, pj) of the double path with everything a pass
    // against it needs (createRef(random=false) + buildBLA, MP.cpp:1750-1758, 4014-4187).
    // the reference at sample (pi, pj, u, v) (createRef, mandel_perturbation.cpp:
    // 4014-4074); with_bla: build its BLA table (else none)
    void set_pixel_reference(int pi, int pj, int u = 0, int v = 0, bool with_bla = true);
    // dc per cell column and row of the (H sub) x (W sub) cell grid relative to the
    // reference's cell (stepParallel, mandel_perturbation.cpp:2300-2312)
    void cell_dc(std::vector<double>& col, std::vector<double>& row) const;
    // ref_pixel_n / ref_pixel_value are final after this (the floatexp path computes
    // an EDE reference pixel on a host thread while the frame's kernels run)
    void ref_pixel_wait();
    // the cutoff-sensitive fallback: run the bounded view-centre reference's brute-force
    // test in prepare() (set by mx_compute, which applies the fallback), and the
    // brute-force value of one output cell (thread-safe)
    bool check_cutoff = false;
    double pixel_value(int i, int j, int u, int v, int& n_out);

  private:
    std::thread ref_job;
    bool job_pending = false;
    int job_n = -1;
    double job_value = -2.0;
    void init_mpf();
    void build_orbit(mpf_srcptr cre, mpf_srcptr cim, int period_len = 0);
    int find_nucleus(mpf_ptr nre, mpf_ptr nim, mpf_srcptr scale, int maxp);
    bool try_periodic(mpf_srcptr scale);
    Fe nucleus_size() const;
    void build_orbit_fixed(mpf_srcptr cre, mpf_srcptr cim);
    void ref_pixel(bool async);
    void sa_step(int i);
    int prepare_dbl(mpf_srcptr scale, std::string& err);
    bool tail_escapes(int extra);
    bool virtual_ref_sensitive();
    double point_value(mpf_srcptr c_re, mpf_srcptr c_im, int mxit, bool ede, int& n_out);
    double point_value_serial(mpf_srcptr c_re, mpf_srcptr c_im, int mxit, bool ede, int& n_out);
    void build_bla(double dcmax);
};

} // namespace mx

struct mx_frame {
    mx::Frame impl;
};

struct mx_ctx {
    int device = 0;
    int cu_count = 0;
    int lds_max = 0;                          // dynamic LDS bytes a workgroup may use
    unsigned char* noalloc_scratch = nullptr; // device address of this device's fixed scratch area
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;     // a second stream, for a pass that runs beside another
    int* cancel_host = nullptr;     // pinned, host-mapped flag
    int* cancel_dev = nullptr;      // device view of the same flag
    int* abort_host = nullptr;      // pinned, host-mapped: stops a pass on the side stream
    int* abort_dev = nullptr;       // device view of the same flag
    std::atomic<bool> busy{false};
    std::string err;
    std::mutex err_mu;
    void set_err(const std::string& e) {
        std::lock_guard<std::mutex> g(err_mu);
        err = e;
    }
};

struct mx_ref {
    mx_ctx* ctx = nullptr;
    MxRefDev dev{};
    int2* lev = nullptr;        // device level table (off, size)
    void* block = nullptr;      // one device allocation for every array
    size_t bytes = 0;
};

struct mx_plan {
    mx_ctx* ctx = nullptr;
    const mx_ref* ref = nullptr;
    MxTileDev tile{};
    int form = MX_FORM_FE64;
    int blocks = 0, threads = 256;
    int minlevel = 0;
    void* block = nullptr;      // dc tables, outputs, queue counters, stats
    size_t bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t n = 0;
    bool ede = false, interior = false;
    MxJob job{};                // floatexp forms: the launch's job (host copy)
    MxJob* job_dev = nullptr;   // and its device copy, inside `block`
    int work_tiles = 0;         // 64-sample work items of one launch
    int grid_request = 0;       // caller's grid size (0: default)
    int nstage = 0;             // plain-double form: orbit entries staged in LDS
    hipStream_t stream = nullptr; // the stream its launch, wait and reads use
};

struct mx_batch {
    mx_ctx* ctx = nullptr;
    std::vector<mx_plan*> plans;
    MxJob* jobs_dev = nullptr;
    int form = MX_FORM_FE64;
    bool ede = false, interior = false;
    int blocks = 0, threads = 256;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct mx_graph {
    mx_ctx* ctx = nullptr;
    std::vector<mx_plan*> plans;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

hipError_t mx_launch_fe64(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream);
hipError_t mx_launch_fe32(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream);
hipError_t mx_launch_sc64(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream);
hipError_t mx_launch_dbl(const MxRefDev& R, const int2* lev, const MxTileDev& T, int blocks, int threads, hipStream_t stream, int nstage);
hipError_t mx_launch_shallow(const MxShallowDev& S, bool ede, int blocks, hipStream_t stream);
