// jh_cli: native driver of libjitterhip for profiling and scripted runs.
//   jh_cli <scene.bin> <warmup> <steps> [copy_mib]
// copy_mib > 0 first copies that many MiB device to device (hipMemcpy), a transfer of known size that
// calibrates HBM traffic counters against the kernels.
// scene.bin (written by jitterphysics2_amd.scenes.save_scene): magic "JHSC", version 1 or 2, then
// int32 counts (bodies, shapes, vertices), float dt, jh_config, jh_body[], jh_shape[], float[3*nv];
// version 2 appends int32 nc and jh_constraint[nc].
// After the warmup the launch shapes are autotuned as in the benchmark.  Prints the mean per-stage
// timings (HIP events inside the library) and the wall time of the timed steps.  A kernel-trace
// profile is cut to the timed steps by their solver launches (the last `steps` of them).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/jitterhip.h"

static int die(const char* what, jh_status s) {
    fprintf(stderr, "%s failed (%d): %s\n", what, s, jh_last_error());
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s scene.bin warmup steps\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) {
        perror("open scene");
        return 2;
    }
    char magic[4];
    int32_t ver, nb, ns, nv;
    float dt;
    jh_config cfg;
    if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "JHSC", 4) != 0 || fread(&ver, 4, 1, f) != 1 ||
        (ver != 1 && ver != 2) ||
        fread(&nb, 4, 1, f) != 1 || fread(&ns, 4, 1, f) != 1 || fread(&nv, 4, 1, f) != 1 || fread(&dt, 4, 1, f) != 1 ||
        fread(&cfg, sizeof(cfg), 1, f) != 1 || nb < 0 || ns < 0 || nv < 0) {
        fprintf(stderr, "bad scene file\n");
        return 2;
    }
    std::vector<jh_body> bodies(nb);
    std::vector<jh_shape> shapes(ns);
    std::vector<float> verts(3 * size_t(nv));
    if (fread(bodies.data(), sizeof(jh_body), nb, f) != size_t(nb) ||
        fread(shapes.data(), sizeof(jh_shape), ns, f) != size_t(ns) ||
        fread(verts.data(), sizeof(float), verts.size(), f) != verts.size()) {
        fprintf(stderr, "truncated scene file\n");
        return 2;
    }
    int32_t nc = 0;
    std::vector<jh_constraint> cons;
    if (ver == 2) {
        if (fread(&nc, 4, 1, f) != 1 || nc < 0) {
            fprintf(stderr, "bad constraint count\n");
            return 2;
        }
        cons.resize(nc);
        if (fread(cons.data(), sizeof(jh_constraint), nc, f) != size_t(nc)) {
            fprintf(stderr, "truncated constraints\n");
            return 2;
        }
    }
    fclose(f);
    int warmup = atoi(argv[2]), steps = atoi(argv[3]);
    int copy_mib = argc > 4 ? atoi(argv[4]) : 0;
    if (copy_mib > 0) {
        size_t bytes = size_t(copy_mib) << 20;
        void *src = nullptr, *dst = nullptr;
        if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess) {
            fprintf(stderr, "copy buffers: hipMalloc failed\n");
            return 1;
        }
        if (hipMemset(src, 1, bytes) != hipSuccess || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            fprintf(stderr, "copy failed\n");
            return 1;
        }
        printf("# copied %zu bytes device to device\n", bytes);
        (void)hipFree(src);
        (void)hipFree(dst);
    }
    jh_world* w = nullptr;
    jh_status s = jh_world_create(&cfg, &w);
    if (s) return die("jh_world_create", s);
    if ((s = jh_world_set_bodies(w, bodies.data(), nb))) return die("jh_world_set_bodies", s);
    if ((s = jh_world_set_shapes(w, shapes.data(), ns, nv ? verts.data() : nullptr, nv)))
        return die("jh_world_set_shapes", s);
    if (nc > 0 && (s = jh_world_set_constraints(w, cons.data(), nc))) return die("jh_world_set_constraints", s);
    for (int i = 0; i < warmup; ++i)
        if ((s = jh_world_step(w, dt))) return die("jh_world_step", s);
    printf("# warmup %d steps done\n", warmup);
    // launch shapes as the benchmark picks them (2 autotune rounds, which advance the simulation)
    jh_tuning tune;
    if (getenv("JH_NO_AUTOTUNE")) {
        printf("# autotune skipped (JH_NO_AUTOTUNE)\n");
    } else {
        if ((s = jh_world_autotune(w, dt, 2, &tune))) return die("jh_world_autotune", s);
        printf("# autotune: solver %d x %d, colour blocks %d, traversal %s (%d leaves per wave)\n", tune.solver,
               tune.solve_blocks, tune.colour_blocks, tune.traversal ? "wavefront" : "per-thread", tune.traversal);
    }
    fflush(stdout);
    double sum[9] = {0};
    jh_stats st;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < steps; ++i) {
        if ((s = jh_world_step(w, dt))) return die("jh_world_step", s);
        jh_world_get_stats(w, &st);
        const double v[9] = {st.ms_broadphase, st.ms_narrowphase, st.ms_epa,    st.ms_contact_cache, st.ms_colouring,
                             st.ms_solve,      st.ms_update,     st.ms_step,   st.ms_solve_kernel};
        for (int k = 0; k < 9; ++k) sum[k] += v[k];
    }
    auto t1 = std::chrono::steady_clock::now();
    double wall = std::chrono::duration<double>(t1 - t0).count();
    for (double& x : sum) x /= steps > 0 ? steps : 1;
    int ndyn = 0;
    for (auto& b : bodies) ndyn += b.motion_type == JH_MOTION_DYNAMIC;
    printf("steps %d wall_s %.6f ms_per_step %.4f body_steps_per_s %.1f pairs %lld arbiters %lld contacts %lld "
           "colours %lld epa %lld | mean ms: broadphase %.3f narrowphase %.3f (epa %.3f) cache %.3f colouring %.3f "
           "solve %.3f (kernel %.3f) update %.3f step %.3f\n",
           steps, wall, 1000.0 * wall / steps, ndyn * steps / wall, (long long)st.pairs, (long long)st.arbiters,
           (long long)st.contacts, (long long)st.colours, (long long)st.epa_calls, sum[0], sum[1], sum[2], sum[3], sum[4],
           sum[5], sum[8], sum[6], sum[7]);
    fflush(stdout);
    jh_world_destroy(w);
    return 0;
}
