// zsc_main.cpp -- zsc_run, the host driver of libzsc: the time loop of the
// reference's main() (src/main.cu:83-667) over the C-ABI, on one GPU.
//
//   zsc_run [-d DIR] [-p f64|f32|bf16|fp16]
//
// Reads DIR/properties.in and DIR/discrete.in (default: the current
// directory), runs relaxStep steps with gravity off and maxStep-1 steps with
// gravity on (src/main.cu:363,496), writes relax_stats.out / stats.out every
// `stat` steps in the format of src/stats.c:73,91 and a runlog.out with the
// sections of src/logsave.c:181-281.  Differences from the reference: the
// statistics are reduced on the GPU in FP64 from the current post-stream
// state (the reference uses moments one step old); VTK output is not written.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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

int main(int argc, char **argv)
{
    std::string dir = ".";
    int prec = ZSC_F32;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-d") && i + 1 < argc) {
            dir = argv[++i];
        } else if (!strcmp(argv[i], "-p") && i + 1 < argc) {
            const char *p = argv[++i];
            prec = !strcmp(p, "f64") ? ZSC_F64 : !strcmp(p, "f32") ? ZSC_F32
                   : !strcmp(p, "bf16") ? ZSC_BF16 : !strcmp(p, "fp16") ? ZSC_FP16 : -1;
            if (prec < 0) {
                fprintf(stderr, "zsc_run: unknown precision %s\n", p);
                return 2;
            }
        } else {
            fprintf(stderr, "usage: zsc_run [-d DIR] [-p f64|f32|bf16|fp16]\n");
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

    printf("*** Multiphase Zheng-Shu-Chew LBM 3D Simulation\n");
    printf("*** MI355X HIP implementation (libzsc)\n\n");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess)
        printf("*** Program running on : %s\n*** Total GPU Memory   : %.4f GB\n\n", prop.name,
               prop.totalGlobalMem / 1073741824.0);

    zsc_ctx *ctx = nullptr;
    CK(zsc_create(in.xmax, in.ymax, in.zmax, prec, nullptr, &ctx));
    CK(zsc_set_params(ctx, &p));
    hipEvent_t e0, e1, e2, e3;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    (void)hipEventCreate(&e3);
    (void)hipEventRecord(e0, 0);
    CK(zsc_state_init_bubbles(ctx, bub.data(), nb));
    printf("Data initialization completed.\n");
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(e1, 0);

    zsc_stats s;
    // relaxation: gravity off (src/main.cu:361-474)
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
    CK(zsc_set_gravity(ctx, 0));
    for (int step = 1; step <= in.relax_step;) {
        const int next = std::min(in.relax_step, (step / in.stat + 1) * in.stat);
        const int chunk = next - step + 1;
        CK(zsc_step(ctx, chunk));
        step += chunk;
        if ((step - 1) % in.stat == 0) {
            CK(zsc_state_stats(ctx, &s));
            stats_line(relax, step - 1, s, inv_relax);
        }
    }
    printf("Relaxation run completed.\n");
    CK(zsc_sync(ctx));
    (void)hipEventRecord(e2, 0);

    // evolution: gravity on (src/main.cu:479-610), maxStep-1 steps
    StatsFile evo{dir + "/stats.out"};
    {
        FILE *fp = fopen(evo.path.c_str(), "w");
        if (fp) {
            fprintf(fp, "#step\t\tux\t\tuy\t\tuz\t\tUmax\t\tVef\t\tRef\n");
            fclose(fp);
        }
    }
    CK(zsc_state_stats(ctx, &s));
    const double inv_init = s.bubble_volume > 0 ? 1.0 / s.bubble_volume : 0.0;
    stats_line(evo, 0, s, inv_init);
    CK(zsc_set_gravity(ctx, 1));
    printf("Starting evolution run ...\n");
    double evo_ms = 0.0;
    const int last = in.max_step - 1;
    for (int step = 1; step <= last;) {
        const int next = std::min(last, (step / in.stat + 1) * in.stat);
        const int chunk = next - step + 1;
        double ms = 0.0;
        CK(zsc_step_timed(ctx, chunk, &ms));
        evo_ms += ms;
        step += chunk;
        if ((step - 1) % in.stat == 0) {
            CK(zsc_state_stats(ctx, &s));
            stats_line(evo, step - 1, s, inv_init);
        }
    }
    printf("Evolution run completed.\n");
    CK(zsc_sync(ctx));
    (void)hipEventRecord(e3, 0);
    (void)hipEventSynchronize(e3);
    float t_setup = 0, t_relax = 0, t_main = 0;
    (void)hipEventElapsedTime(&t_setup, e0, e1);
    (void)hipEventElapsedTime(&t_relax, e1, e2);
    (void)hipEventElapsedTime(&t_main, e2, e3);
    CK(zsc_state_stats(ctx, &s));
    const double cells = (double)in.xmax * in.ymax * in.zmax;
    const double mlups = last > 0 && evo_ms > 0 ? cells * last / (evo_ms * 1e3) : 0.0;
    printf("Setup time      : %6.3f ms\n", t_setup);
    printf("Relaxation time : %6.3f ms\n", t_relax);
    printf("Evolution time  : %6.3f ms\n", t_main);
    printf("Time for lb step: %6.3f ms\n", last > 0 ? evo_ms / last : 0.0);
    printf("MLUPS (step kernels only): %.1f\n", mlups);

    FILE *fp = fopen((dir + "/runlog.out").c_str(), "w");
    if (fp) {
        fprintf(fp, "***\n*** Multiphase Zheng-Shu-Chew LBM 3D Simulation\n");
        fprintf(fp, "*** MI355X HIP implementation (libzsc)\n***\n\n");
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
        fprintf(fp, "LBM step Time   = %9.6f\n", last > 0 ? evo_ms * 1e-3 / last : 0.0);
        fprintf(fp, "MLUPS           = %9.6f\n", mlups);
        fclose(fp);
    }
    zsc_destroy(ctx);
    return 0;
}
