// cli.cpp -- `GFSM <query.g> <data.g>`: the reference's command line
// (main.cu:56-63) over libgfsm.  Prints the reference's CSV row
// (mainscripts/runLabel.sh:19 column order) and exits with its codes
// (232/233/234 for input errors, 1 for query errors).
// --cache PATH keeps a binary image of the data graph: a later run whose
// image is newer than the data file and of the same mode loads it instead of
// parsing the text and building the layout again.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "gfsm.h"

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static std::string base_name(const std::string &p)
{
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

// Data graph through the binary image at `cache` when that image is usable,
// else from the text (writing the image for next time; a failed write only
// warns).
static gfsm_status load_data(const char *path, const char *cache, uint32_t mode, gfsm_graph **d)
{
    if (cache) {
        struct stat ts, cs;
        auto ns = [](const struct stat &x) { return (long long)x.st_mtim.tv_sec * 1000000000ll + x.st_mtim.tv_nsec; };
        if (stat(cache, &cs) == 0 && stat(path, &ts) == 0 && ns(cs) >= ns(ts) &&
            gfsm_graph_load_bin(cache, d) == GFSM_OK) {
            uint32_t nv, md, m;
            uint64_t ne;
            if (gfsm_graph_info(*d, &nv, &ne, &md, &m) == GFSM_OK && m == mode) return GFSM_OK;
            gfsm_graph_free(*d);
            *d = nullptr;
        }
    }
    gfsm_status st = gfsm_graph_load_tve(path, mode, d);
    if (st == GFSM_OK && cache && gfsm_graph_save_bin(*d, cache) != GFSM_OK)
        std::fprintf(stderr, "warning: could not write %s: %s\n", cache, gfsm_last_error());
    return st;
}

// The first `limit` embeddings on stderr, one per line: "q0->v q1->w ...";
// from the result when it holds rows, else read back from output.fs.
static void print_embeddings(const gfsm_result *r, const char *out, uint64_t limit)
{
    const uint32_t *rows = nullptr;
    uint32_t w = 0;
    uint64_t n = 0;
    gfsm_result_rows(r, &rows, &w, &n);
    std::vector<uint32_t> buf;
    if (!n && out) {
        FILE *f = std::fopen(out, "rb");
        uint32_t hdr[2];
        if (f && std::fread(hdr, sizeof(uint32_t), 2, f) == 2) {
            w = hdr[1];
            n = std::min<uint64_t>(hdr[0], limit);
            buf.resize(n * w);
            n = w ? std::fread(buf.data(), sizeof(uint32_t) * w, n, f) : 0;
            rows = buf.data();
        }
        if (f) std::fclose(f);
    }
    for (uint64_t i = 0; i < n && i < limit; i++) {
        for (uint32_t c = 0; c < w; c++) std::fprintf(stderr, "%s%u->%u", c ? " " : "", c, rows[i * w + c]);
        std::fprintf(stderr, "\n");
    }
}

static int die(gfsm_status st)
{
    std::fprintf(stderr, "\nGFSM error %d: %s\n", st, gfsm_last_error());
    return st;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr,
                     "usage: %s <query.g> <data.g> [--mode ref|vlabel] [--out output.fs] [--sorted]\n"
                     "          [--mem-budget BYTES] [--device N] [--digest] [--cache data.bin] [--stream] [--print N]\n",
                     argv[0]);
        return 2;
    }
    uint32_t mode = GFSM_MODE_REF;
    const char *out = nullptr, *cache = nullptr;
    uint64_t print_rows = 0;
    bool stream = false;
    uint64_t budget = 0;
    int device = 0;
    uint32_t flags = 0;
    for (int i = 3; i < argc; i++) {
        if (!std::strcmp(argv[i], "--mode") && i + 1 < argc) {
            mode = !std::strcmp(argv[++i], "vlabel") ? GFSM_MODE_VLABEL : GFSM_MODE_REF;
        } else if (!std::strcmp(argv[i], "--out") && i + 1 < argc) {
            out = argv[++i];
        } else if (!std::strcmp(argv[i], "--mem-budget") && i + 1 < argc) {
            budget = std::strtoull(argv[++i], nullptr, 10);
        } else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else if (!std::strcmp(argv[i], "--stream")) {
            stream = true;
        } else if (!std::strcmp(argv[i], "--print") && i + 1 < argc) {
            print_rows = std::strtoull(argv[++i], nullptr, 10);
        } else if (!std::strcmp(argv[i], "--cache") && i + 1 < argc) {
            cache = argv[++i];
        } else if (!std::strcmp(argv[i], "--digest")) {
            flags |= GFSM_DIGEST;
        } else if (!std::strcmp(argv[i], "--sorted")) {
            flags |= GFSM_SORTED;
        }
    }
    if (out || print_rows) flags |= GFSM_ROWS;
    std::printf("\n%s,%s,", base_name(argv[1]).c_str(), base_name(argv[2]).c_str());
    const double t_start = now_s();
    // Inputs are parsed and the query compiled before the device is opened,
    // so malformed input reports 232/233/234/1 on any machine.
    gfsm_graph *q = nullptr, *d = nullptr;
    double t = now_s();
    gfsm_status st = gfsm_graph_load_tve(argv[1], mode, &q);
    if (st) return die(st);
    const double q_parse = now_s() - t;
    t = now_s();
    st = load_data(argv[2], cache, mode, &d);
    if (st) return die(st);
    const double d_parse = now_s() - t;
    uint32_t qn = 0;
    st = gfsm_plan_query(q, &qn, nullptr, nullptr, nullptr, nullptr);
    if (st) return die(st);

    gfsm_ctx *ctx = nullptr;
    t = now_s();
    st = gfsm_open(device, budget, &ctx);
    if (st) return die(st);
    std::printf("%fs,", now_s() - t);  // Preallocate Time
    std::printf("%fs,", q_parse);      // Query Parse Time
    std::printf("%fs,", d_parse);      // Data Parse Time

    gfsm_match_opts o{};
    o.flags = flags;
    gfsm_result *r = nullptr;
    const double t_gpu = now_s();
    // --stream: rows go to output.fs chunk by chunk during the match (the
    // write is then inside GPU Time, and Solution Write Time is ~0).
    st = (out && stream) ? gfsm_match_write_fs(ctx, d, q, &o, out, &r) : gfsm_match(ctx, d, q, &o, &r);
    if (st) return die(st);
    const double gpu_s = now_s() - t_gpu;
    gfsm_stats s;
    gfsm_result_stats(r, &s);
    uint64_t count = 0;
    gfsm_result_count(r, &count);
    std::printf("%fs,", s.device_ms / 1e3);  // Kernel Time
    std::printf("n/a,n/a,");                  // Reallocate / ReKernel: no retry exists
    std::printf("%llu,", (unsigned long long)count);
    t = now_s();
    if (out && !stream) {
        st = gfsm_result_write_fs(r, out);
        if (st) return die(st);
    }
    std::printf("%fs,", now_s() - t);  // Solution Write Time
    std::printf("%fs,", gpu_s);        // GPU Time
    if (flags & GFSM_DIGEST) {
        uint64_t dg[2];
        gfsm_result_digest(r, dg);
        std::fprintf(stderr, "digest=%016llx%016llx\n", (unsigned long long)dg[0], (unsigned long long)dg[1]);
    }
    if (print_rows) print_embeddings(r, out, print_rows);
    gfsm_result_free(r);
    gfsm_graph_free(q);
    gfsm_graph_free(d);
    gfsm_close(ctx);
    std::printf("%fs,", now_s() - t_start);  // End to End Time
    std::printf("\n");
    return 0;
}
