/*
 * oracle_cli.c -- command-line front end of the CPU oracle (test
 * infrastructure only).  Usage mirrors the reference CLI (main.cu:56-63):
 *   gfsm_oracle <query.g> <data.g> [--mode ref|vlabel] [--threads N] [--out output.fs]
 * Prints count, digest and wall time; exit codes are the reference's.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <query.g> <data.g> [--mode ref|vlabel] [--threads N] [--out file]\n", argv[0]);
        return 2;
    }
    uint32_t mode = ORC_MODE_REF;
    int threads = 0;
    const char *outp = NULL;
    for (int i = 3; i < argc; i++) {
        if (!strcmp(argv[i], "--mode") && i + 1 < argc) {
            mode = !strcmp(argv[++i], "vlabel") ? ORC_MODE_VLABEL : ORC_MODE_REF;
        } else if (!strcmp(argv[i], "--threads") && i + 1 < argc) {
            threads = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--out") && i + 1 < argc) {
            outp = argv[++i];
        }
    }
    orc_edges *q = NULL, *d = NULL;
    double t0 = now_s();
    int rc = orc_load_tve(argv[1], &q);
    if (rc) return rc;
    rc = orc_load_tve(argv[2], &d);
    if (rc) return rc;
    double t1 = now_s();
    orc_result *r = NULL;
    rc = orc_solve(d, q, mode, 0, 0xFFFFFFFFu, threads, outp ? (1ull << 32) - 1 : 0, &r);
    double t2 = now_s();
    if (rc) return rc;
    printf("count=%llu digest=%016llx%016llx parse_s=%.6f match_s=%.6f\n",
           (unsigned long long)r->count, (unsigned long long)r->digest[0],
           (unsigned long long)r->digest[1], t1 - t0, t2 - t1);
    if (outp) rc = orc_fs_write(outp, r->rows, r->nrows, r->width);
    orc_result_free(r);
    orc_edges_free(q);
    orc_edges_free(d);
    return rc;
}
