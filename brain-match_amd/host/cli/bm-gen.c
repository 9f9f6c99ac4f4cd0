/*
 * bm-gen -- write a seeded synthetic graph pair in the reference CLIs'
 * formats.
 *
 *   bm-gen powerlaw <seed> <n> <m> <outdir>   (plain refinement workload)
 *   bm-gen chunglu  <seed> <n> <m> <outdir>   (full-size sweep workload)
 *
 * Writes male.csv / female.csv ("Source,Target,Weight"), init_mf.csv and
 * truth_mf.csv ("Male Node ID,Female Node ID", m<id>,f<id>), init_plain.csv
 * ("Male ID,Female ID", <id>,<id>) and the ordering files order_male.csv /
 * order_female.csv that graph-alignment-optimized requires.
 */
#include <sys/stat.h>

#include "cli_common.h"

static void path(char *buf, size_t sz, const char *dir, const char *name) { snprintf(buf, sz, "%s/%s", dir, name); }

int main(int argc, char *argv[]) {
    if (argc != 6) {
        fprintf(stderr, "Usage: %s <powerlaw|chunglu> <seed> <n> <m> <outdir>\n", argv[0]);
        return 1;
    }
    uint64_t seed = strtoull(argv[2], NULL, 10);
    int32_t n = atoi(argv[3]);
    int64_t m = atoll(argv[4]);
    const char *dir = argv[5];
    mkdir(dir, 0755);
    bm_instance in;
    int rc;
    if (strcmp(argv[1], "powerlaw") == 0) rc = bm_gen_powerlaw(seed, n, m, 0.10, 1.0, &in);
    else if (strcmp(argv[1], "chunglu") == 0) rc = bm_gen_chunglu(seed, n, m, 0.15, 0.15, 0.30, &in);
    else rc = BM_EINVAL;
    if (rc) {
        fprintf(stderr, "generation failed: %s\n", bm_strerror(rc));
        return 1;
    }
    char p[4096];
    path(p, sizeof p, dir, "male.csv");
    bm_edges_write_csv(p, "Source,Target,Weight", in.s1, in.d1, in.w1, in.e1);
    path(p, sizeof p, dir, "female.csv");
    bm_edges_write_csv(p, "Source,Target,Weight", in.s2, in.d2, in.w2, in.e2);
    path(p, sizeof p, dir, "init_mf.csv");
    bm_mapping_write_mf(p, in.p0, in.n);
    path(p, sizeof p, dir, "truth_mf.csv");
    bm_mapping_write_mf(p, in.pi, in.n);
    path(p, sizeof p, dir, "init_plain.csv");
    FILE *f = fopen(p, "w");
    fprintf(f, "Male ID,Female ID\n");
    for (int32_t i = 1; i <= in.n; i++) fprintf(f, "%d,%d\n", i, in.p0[i]);
    fclose(f);
    const char *ords[2] = {"order_male.csv", "order_female.csv"};
    for (int j = 0; j < 2; j++) {
        path(p, sizeof p, dir, ords[j]);
        f = fopen(p, "w");
        fprintf(f, "node,order\n");
        for (int32_t i = 1; i <= in.n; i++) fprintf(f, "%d,%d\n", i, i);
        fclose(f);
    }
    printf("n=%d e1=%lld e2=%lld -> %s\n", in.n, (long long)in.e1, (long long)in.e2, dir);
    bm_instance_free(&in);
    return 0;
}
