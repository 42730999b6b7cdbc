// zsc_main.cpp -- zsc_run, the host driver of libzsc: the time loop of the
// reference's main() (src/main.cu:83-667) over the C-ABI, on one GPU.
//
//   zsc_run [-d DIR] [-p f64|f32|f32d|bf16|fp16] [-v ascii|binary|none]
//           [-r RESTART_IN] [-s RESTART_OUT] [-t TRIALS]
//
// Reads DIR/properties.in and DIR/discrete.in (default: the current
// directory), runs relaxStep steps with gravity off and maxStep-1 steps with
// gravity on (src/main.cu:363,496), writes relax_stats.out / stats.out every
// `stat` steps in the format of src/stats.c:73,91, VTK files every `save`
// steps named as src/vtksave.c:39-40 (relax_data_00_NNNNNNN.vtk,
// data_00_NNNNNNN.vtk), and a runlog.out with the sections of
// src/logsave.c:181-281, including the Laplace-law check after relaxation
// (src/logsave.c:155-162).  -r resumes the evolution phase from a restart
// file (no relaxation); -s writes one at the end.  -t keeps the fastest of
// TRIALS lattice allocations (zsc_place_lattices; default 4, 0 for one
// allocation): the step's time depends on the physical pages behind the
// lattices.  -p f32d stores g as FP32 deviations from w_i * rho_mean (the
// precision to use for large density ratios).  Differences from the
// reference: the statistics are reduced on the GPU in FP64 from the current
// post-stream state (the reference uses moments one step old).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/zsc.h"

static int die(const char *what, int code)
{
    fprintf(stderr, "zsc_run: %s failed (%d): %s\n", what, code, zsc_last_error());
    return 1;
}

#define CK(call)                         \
    do {                                 \
        int rc_ = (call);                \
        if (rc_ != ZSC_OK) return die(#call, rc_); \
    } while (0)

struct StatsFile {
    std::string path;
    double inv_vol = 0.0;
};

static const double kInvPi = 0.31830988618379067154;

// "HH:MM on YYYY-MM-DD", the stamp of src/logsave.c:194-199,272-277
static std::string stamp()
{
    char b[32] = "";
    time_t t = time(nullptr);
    struct tm tmv;
    if (localtime_r(&t, &tmv)) strftime(b, sizeof b, "%H:%M on %Y-%m-%d", &tmv);
    return b;
}

// Values a restart file carries for this driver (zsc_state_save_ex): the
// bubble volume at the start of evolution, which every later Vef divides by
// (src/stats.c:66,84), and the relaxation results of src/logsave.c:155-167.
enum { kUserInitVol, kUserRelaxVef, kUserPdif, kUserPerr, kUserRef, kUserUmax, kUserCount };

static void stats_line(const StatsFile &sf, int step, const zsc_stats &s, double inv_vol)
{
    FILE *fp = fopen(sf.path.c_str(), "a");
    if (!fp) return;
    const double vol = s.bubble_volume;
    const double iv = vol > 0 ? 1.0 / vol : 0.0;
    const double ref = pow(vol * kInvPi * 0.75, 1.0 / 3.0);
    fprintf(fp, "%8d\t%e\t%e\t%e\t%e\t%e\t%e\n", step, s.ux_sum * iv, s.uy_sum * iv, s.uz_sum * iv,
            sqrt(s.umax_sq), vol * inv_vol, ref);
    fclose(fp);
}

// VTK file of src/vtksave.c:39-40 for one rank (vproc = 0).
static int vtk(zsc_ctx *ctx, const std::string &dir, int mode, bool relax, int step)
{
    if (mode < 0) return ZSC_OK;
    char name[64];
    snprintf(name, sizeof name, relax ? "relax_data_%02d_%07d.vtk" : "data_%02d_%07d.vtk", 0, step);
    return zsc_vtk_save(ctx, (dir + "/" + name).c_str(), mode);
}

// Runs steps first..last (inclusive) in chunks that end on the stat and save
// intervals, with stats lines and VTK files there.
static int run_phase(zsc_ctx *ctx, const zsc_inputs &in, const std::string &dir, int vmode,
                     bool relax, const StatsFile &sf, double inv_vol, int first, int last,
                     double *kernel_ms)
{
    for (int step = first; step <= last;) {
        // the first multiple of stat or save at or after this step
        int next = std::min(last, (step + in.stat - 1) / in.stat * in.stat);
        next = std::min(next, (step + in.save - 1) / in.save * in.save);
        const int chunk = next - step + 1;
        double ms = 0.0;
        if (int r = zsc_step_timed(ctx, chunk, &ms)) return die("zsc_step", r);
        if (kernel_ms) *kernel_ms += ms;
        step += chunk;
        const int done = step - 1;
        if (done % in.stat == 0) {
            zsc_stats s;
            if (int r = zsc_state_stats(ctx, &s)) return die("zsc_state_stats", r);
            stats_line(sf, done, s, inv_vol);
        }
        if (done % in.save == 0)
            if (int r = vtk(ctx, dir, vmode, relax, done)) return die("zsc_vtk_save", r);
    }
    return 0;
}

int main(int argc, char **argv)
{
    std::string dir = ".", restart_in, restart_out;
    int prec = ZSC_F32;
    int vmode = 0;  // 0 ASCII (as the reference), 1 binary, -1 none
    int trials = 4;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-d") && i + 1 < argc) {
            dir = argv[++i];
        } else if (!strcmp(argv[i], "-r") && i + 1 < argc) {
            restart_in = argv[++i];
        } else if (!strcmp(argv[i], "-s") && i + 1 < argc) {
            restart_out = argv[++i];
        } else if (!strcmp(argv[i], "-t") && i + 1 < argc) {
            trials = atoi(argv[++i]);
            if (trials < 0 || trials > 16) {
                fprintf(stderr, "zsc_run: -t takes 0..16 trials\n");
                return 2;
            }
        } else if (!strcmp(argv[i], "-v") && i + 1 < argc) {
            const char *v = argv[++i];
            vmode = !strcmp(v, "ascii") ? 0 : !strcmp(v, "binary") ? 1 : !strcmp(v, "none") ? -1 : -2;
            if (vmode == -2) {
                fprintf(stderr, "zsc_run: unknown VTK mode %s\n", v);
                return 2;
            }
        } else if (!strcmp(argv[i], "-p") && i + 1 < argc) {
            const char *p = argv[++i];
            prec = !strcmp(p, "f64") ? ZSC_F64 : !strcmp(p, "f32") ? ZSC_F32
                   : !strcmp(p, "f32d") ? ZSC_F32D : !strcmp(p, "bf16") ? ZSC_BF16
                   : !strcmp(p, "fp16") ? ZSC_FP16 : -1;
            if (prec < 0) {
                fprintf(stderr, "zsc_run: unknown precision %s\n", p);
                return 2;
            }
        } else {
            fprintf(stderr, "usage: zsc_run [-d DIR] [-p f64|f32|f32d|bf16|fp16] [-v ascii|binary|none] "
                            "[-r RESTART_IN] [-s RESTART_OUT] [-t TRIALS]\n");
            return 2;
        }
    }
    zsc_inputs in;
    CK(zsc_read_inputs((dir + "/properties.in").c_str(), &in));
    int nb = 0;
    CK(zsc_read_bubbles((dir + "/discrete.in").c_str(), nullptr, 0, &nb));
    std::vector<double> bub((size_t)(nb > 0 ? nb : 1) * 4);
    CK(zsc_read_bubbles((dir + "/discrete.in").c_str(), bub.data(), nb, &nb));
    zsc_params p;
    CK(zsc_derive_params(&in, nb > 0 ? bub[3] : 0.0, &p));
    if (in.stat <= 0) in.stat = 1 << 30;
    if (in.save <= 0) in.save = 1 << 30;

    const std::string started = stamp();
    printf("*** Multiphase Zheng-Shu-Chew LBM 3D Simulation\n");
    printf("*** MI355X HIP implementation (libzsc)\n\n");
    hipDeviceProp_t prop;
    memset(&prop, 0, sizeof prop);
    const bool have_prop = hipGetDeviceProperties(&prop, 0) == hipSuccess;
    if (have_prop)
        printf("*** Program running on : %s\n*** Total GPU Memory   : %.4f GB\n\n", prop.name,
               prop.totalGlobalMem / 1073741824.0);

    zsc_ctx *ctx = nullptr;
    CK(zsc_create(in.xmax, in.ymax, in.zmax, prec, nullptr, &ctx));
    CK(zsc_set_params(ctx, &p));
    if (trials > 0) {
        double ms[16];
        CK(zsc_place_lattices(ctx, trials, ms));
        printf("Lattice placement: %d trials, ms per step", trials);
        for (int t = 0; t < trials; ++t) printf(" %.4f", ms[t]);
        printf("\n\n");
    }
    hipEvent_t e0, e1, e2, e3;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    (void)hipEventCreate(&e3);
    (void)hipEventRecord(e0, 0);
    zsc_stats s;
    int64_t resumed = -1;
    double laplace[4] = {0, 0, 0, 0};  // Pdif, Perr, Ref, Umax after relaxation
    double relax_vef = 0.0;
    double user[kUserCount] = {0, 0, 0, 0, 0, 0};
    int n_user = 0;
    if (!restart_in.empty()) {
        CK(zsc_state_load_ex(ctx, restart_in.c_str(), &resumed, user, kUserCount, &n_user));
        printf("Restarted from %s (step %lld).\n", restart_in.c_str(), (long long)resumed);
        if (n_user < kUserCount)
            fprintf(stderr, "zsc_run: %s carries no reference bubble volume; Vef is relative to "
                            "the resumed state\n", restart_in.c_str());
    } else {
        CK(zsc_state_init_bubbles(ctx, bub.data(), nb));
    }
    printf("Data initialization completed.\n");
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(e1, 0);

    if (resumed < 0) {
        // relaxation: gravity off (src/main.cu:296-474)
        StatsFile relax{dir + "/relax_stats.out"};
        {
            FILE *fp = fopen(relax.path.c_str(), "w");
            if (fp) {
                fprintf(fp, "#step\t\tux\t\tuy\t\tuz\t\tUmax\t\tVef\t\tRef\n");
                fclose(fp);
            }
        }
        CK(zsc_state_stats(ctx, &s));
        const double inv_relax = s.bubble_volume > 0 ? 1.0 / s.bubble_volume : 0.0;
        stats_line(relax, 0, s, inv_relax);
        CK(vtk(ctx, dir, vmode, true, 0));
        CK(zsc_set_gravity(ctx, 0));
        if (int r = run_phase(ctx, in, dir, vmode, true, relax, inv_relax, 1, in.relax_step, nullptr))
            return r;
        // Laplace-law check of src/logsave.c:155-162 (single static bubble)
        CK(zsc_state_stats(ctx, &s));
        const double cells = (double)in.xmax * in.ymax * in.zmax;
        const double vin = s.bubble_volume, vout = cells - vin;
        const double pin = vin > 0 ? s.p_in_sum / vin : 0.0, pout = vout > 0 ? s.p_out_sum / vout : 0.0;
        const double ro = nb > 0 ? bub[3] : 0.0;
        laplace[0] = pin - pout;
        laplace[1] = ro > 0 ? 0.5 * (2.0 * in.sigma / ro - laplace[0]) * ro / in.sigma : 0.0;
        laplace[2] = pow(vin * kInvPi * 0.75, 1.0 / 3.0);
        laplace[3] = sqrt(s.umax_sq);
        relax_vef = vin * inv_relax;
        printf("Relaxation run completed.\n");
    }
    CK(zsc_sync(ctx));
    (void)hipEventRecord(e2, 0);

    // evolution: gravity on (src/main.cu:476-610), maxStep-1 steps
    StatsFile evo{dir + "/stats.out"};
    const int first = resumed < 0 ? 1 : (int)resumed + 1;
    {
        FILE *fp = fopen(evo.path.c_str(), resumed < 0 ? "w" : "a");
        if (fp) {
            if (resumed < 0) fprintf(fp, "#step\t\tux\t\tuy\t\tuz\t\tUmax\t\tVef\t\tRef\n");
            fclose(fp);
        }
    }
    CK(zsc_state_stats(ctx, &s));
    // the reference volume of every evolution Vef: this state's, or on a
    // resumed run the one the original run started its evolution with
    const double init_vol = resumed >= 0 && n_user >= kUserCount ? user[kUserInitVol] : s.bubble_volume;
    const double inv_init = init_vol > 0 ? 1.0 / init_vol : 0.0;
    if (resumed < 0) {
        user[kUserInitVol] = init_vol;
        user[kUserRelaxVef] = relax_vef;
        for (int i = 0; i < 4; ++i) user[kUserPdif + i] = laplace[i];
    } else if (n_user >= kUserCount) {
        relax_vef = user[kUserRelaxVef];
        for (int i = 0; i < 4; ++i) laplace[i] = user[kUserPdif + i];
    }
    if (resumed < 0) {
        stats_line(evo, 0, s, inv_init);
        CK(vtk(ctx, dir, vmode, false, 0));
    }
    CK(zsc_set_gravity(ctx, 1));
    printf("Starting evolution run ...\n");
    double evo_ms = 0.0;
    const int last = in.max_step - 1;
    if (int r = run_phase(ctx, in, dir, vmode, false, evo, inv_init, first, last, &evo_ms)) return r;
    printf("Evolution run completed.\n");
    CK(zsc_sync(ctx));
    (void)hipEventRecord(e3, 0);
    (void)hipEventSynchronize(e3);
    float t_setup = 0, t_relax = 0, t_main = 0;
    (void)hipEventElapsedTime(&t_setup, e0, e1);
    (void)hipEventElapsedTime(&t_relax, e1, e2);
    (void)hipEventElapsedTime(&t_main, e2, e3);
    CK(zsc_state_stats(ctx, &s));
    // final configuration, named with the loop counter's end value maxStep
    // as src/main.cu:628-634 does
    CK(vtk(ctx, dir, vmode, false, in.max_step));
    if (!restart_out.empty())
        CK(zsc_state_save_ex(ctx, restart_out.c_str(), last, user, kUserCount));
    const double cells = (double)in.xmax * in.ymax * in.zmax;
    const int nrun = last >= first ? last - first + 1 : 0;
    const double mlups = nrun > 0 && evo_ms > 0 ? cells * nrun / (evo_ms * 1e3) : 0.0;
    // the reference's own figures (src/main.cu:624, src/logsave.c:262-263):
    // the main-loop time, stats and VTK output included, divided by maxStep
    // although maxStep-1 steps ran, over a grid that counts the two z ghost
    // planes of src/main.cu:161-164
    const double ref_cells = (double)in.xmax * in.ymax * (in.zmax + 2.0);
    const double ref_step_ms = in.max_step > 0 ? t_main / in.max_step : 0.0;
    const double ref_mlups = t_main > 0 ? in.max_step * ref_cells * 1e-6 / (t_main * 1e-3) : 0.0;
    printf("Setup time      : %6.3f ms\n", t_setup);
    printf("Relaxation time : %6.3f ms\n", t_relax);
    printf("Evolution time  : %6.3f ms\n", t_main);
    printf("Time for lb step: %6.3f ms (reference formula: main loop / maxStep)\n", ref_step_ms);
    printf("MLUPS           : %.1f (reference formula: maxStep * nx*ny*(nz+2) / main loop)\n", ref_mlups);
    printf("Step kernels    : %6.3f ms per step, %.1f MLUPS (nx*ny*nz * %d steps / kernel time)\n",
           nrun > 0 ? evo_ms / nrun : 0.0, mlups, nrun);

    FILE *fp = fopen((dir + "/runlog.out").c_str(), "w");
    if (fp) {
        fprintf(fp, "***\n*** Multiphase Zheng-Shu-Chew LBM 3D Simulation\n");
        fprintf(fp, "*** MI355X HIP implementation (libzsc)\n***\n");
        if (have_prop) {
            const double gb = 1073741824.0, mb = 1048576.0, dm = (double)prop.totalGlobalMem;
            fprintf(fp, "*** Program running on : %s\n", prop.name);
            if (dm >= gb)
                fprintf(fp, "*** Total GPU Memory   : %.4f GB\n", dm / gb);
            else
                fprintf(fp, "*** Total GPU Memory   : %.4f MB\n", dm / mb);
            fprintf(fp, "***\n");
        }
        fprintf(fp, "*** Simulation started at %s\n***\n\n", started.c_str());
        fprintf(fp, "INPUT PARAMETERS\n");
        fprintf(fp, "Evolution Steps    = %d\n", in.max_step);
        fprintf(fp, "Relaxation Steps   = %d\n", in.relax_step);
        fprintf(fp, "Domain Size in X   = %d\n", in.xmax);
        fprintf(fp, "Domain Size in Y   = %d\n", in.ymax);
        fprintf(fp, "Domain Size in Z   = %d\n", in.zmax);
        fprintf(fp, "Interface Width    = %6.3f\n", in.width);
        fprintf(fp, "Interface Tension  = %6.3f\n", in.sigma);
        fprintf(fp, "Interface Mobility = %6.3f\n", in.gamma);
        fprintf(fp, "Low Fluid Density  = %6.3f\n", in.rho_l);
        fprintf(fp, "High Fluid Density = %6.3f\n", in.rho_h);
        fprintf(fp, "Tau (Density)      = %6.3f\n", in.tau_rho);
        fprintf(fp, "Tau (Phase)        = %6.3f\n\n", in.tau_phi);
        // launch configuration (the reference's CUDA CONFIGURATION section
        // prints the properties.in block size its 2-D grid uses; here that
        // value is read and printed, and the fused kernel's tile follows)
        int ty = 0, zc = 0, fl = 0;
        (void)zsc_get_fused_config(ctx, &ty, &zc, &fl);
        const bool wide = (fl & 16) && prec != ZSC_F64 && in.xmax % 128 == 0 && (ty == 4 || ty == 8);
        const int tile_x = wide ? 128 : 64;
        fprintf(fp, "GPU CONFIGURATION\n");
        fprintf(fp, "GPUs used    = %d\n", 1);
        fprintf(fp, "BLOCK_SIZE_X = %d\n", in.block_x);
        fprintf(fp, "BLOCK_SIZE_Y = %d\n", in.block_y);
        fprintf(fp, "Fused tile   = %d x %d cells, %d-plane z chunks, flags %d\n", tile_x, ty,
                zc > 0 ? zc : in.zmax, fl);
        fprintf(fp, "Precision    = %s\n\n",
                prec == ZSC_F64 ? "f64" : prec == ZSC_F32 ? "f32" : prec == ZSC_F32D ? "f32d"
                : prec == ZSC_BF16 ? "bf16" : "fp16");
        // device memory the context holds (two lattices, five moment arrays
        // at the compute precision, an FP64 staging array) and the driver's
        // host buffers (bubbles; VTK output streams through 2^22-cell slabs)
        {
            const double gb = 1073741824.0, mb = 1048576.0;
            const double rsize = prec == ZSC_F64 ? 8.0 : 4.0;
            const double dist = zsc_bytes_per_cell(prec) * cells;
            const double aux = 5.0 * rsize * cells + 8.0 * cells;
            const double cpu = (double)bub.size() * sizeof(double) +
                               (vmode >= 0 ? 3.0 * (rsize + 36.0) * (double)(1 << 22) : 0.0);
            const double gpu = dist + aux;
            const double u = gpu < gb ? mb : gb;
            fprintf(fp, "MEMORY USAGE (%s)\n", gpu < gb ? "Mb" : "Gb");
            fprintf(fp, "Distributions         = %4.2f\n", dist / u);
            fprintf(fp, "Auxiliary Arrays      = %4.2f\n", aux / u);
            fprintf(fp, "Total CPU Memory Used = %4.2f\n", cpu / u);
            fprintf(fp, "Total GPU Memory Used = %4.2f\n\n", gpu / u);
        }
        if (resumed >= 0) fprintf(fp, "RESTART\nResumed at step   = %lld\n\n", (long long)resumed);
        if (resumed < 0 || n_user >= kUserCount) {
            fprintf(fp, "RELAXATION RESULTS\n");
            fprintf(fp, "Effective Radius   = %e\n", laplace[2]);
            fprintf(fp, "Phase Conservation = %e\n", relax_vef);
            fprintf(fp, "(Pin - Pout)       = %e\n", laplace[0]);
            fprintf(fp, "Laplace Error      = %e\n", laplace[1]);
            fprintf(fp, "Parasitic Velocity = %e\n\n", laplace[3]);
        }
        const double vol = s.bubble_volume;
        fprintf(fp, "OUTPUT RESULTS\n");
        fprintf(fp, "Effective Radius   = %e\n", pow(vol * kInvPi * 0.75, 1.0 / 3.0));
        fprintf(fp, "Phase Conservation = %e\n", vol * inv_init);
        fprintf(fp, "Maximum Velocity   = %e\n", sqrt(s.umax_sq));
        fprintf(fp, "Total phi (FP64)   = %.17e\n", s.sum_phi);
        fprintf(fp, "Total rho (FP64)   = %.17e\n\n", s.sum_rho);
        fprintf(fp, "TIMING INFORMATION (seconds)\n");
        fprintf(fp, "Setup Time      = %9.6f\n", t_setup * 1e-3);
        fprintf(fp, "Relaxation Time = %9.6f\n", t_relax * 1e-3);
        fprintf(fp, "Main Loop Time  = %9.6f\n", t_main * 1e-3);
        fprintf(fp, "LBM step Time   = %9.6f\n", ref_step_ms * 1e-3);
        fprintf(fp, "MLUPS           = %9.6f\n\n", ref_mlups);
        fprintf(fp, "STEP KERNELS (nx*ny*nz cells, %d steps, kernel time only)\n", nrun);
        fprintf(fp, "LBM step Time   = %9.6f\n", nrun > 0 ? evo_ms * 1e-3 / nrun : 0.0);
        fprintf(fp, "MLUPS           = %9.6f\n\n", mlups);
        fprintf(fp, "***\n*** Simulation completed at %s\n***\n", stamp().c_str());
        fclose(fp);
    }
    zsc_destroy(ctx);
    return 0;
}
