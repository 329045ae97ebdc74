// Host sanitizer run of libgfsm's untrusted-input code: the t/v/e parser
// (graph.cpp), the binary image loader (graphio.cpp), the layout builder and
// the planner (plan.cpp), built with AddressSanitizer and UBSan (no device
// code is involved).  Exercises valid inputs, malformed text files, damaged
// and crafted images, and plans of generated queries; exits non-zero when a
// status differs from what the input calls for.  Usage: asan_main <tmpdir>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host.h"

using namespace gfsm;

static int failures = 0;

static void expect(bool ok, const std::string &what)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        failures++;
    }
}

static void write_file(const std::string &path, const std::string &text)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
}

static std::string read_file(const std::string &path)
{
    std::string s;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return s;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return s;
}

static gfsm_status load_text(const std::string &dir, const std::string &name, const std::string &text,
                             uint32_t mode = GFSM_MODE_REF)
{
    const std::string p = dir + "/" + name + ".g";
    write_file(p, text);
    HostGraph *g = nullptr;
    Status st = load_tve(p.c_str(), mode, g);
    delete g;
    return st.code;
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    // --- the t/v/e parser on valid and malformed text
    expect(load_text(dir, "readme", "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1") == GFSM_OK, "readme");
    expect(load_text(dir, "vl", "t # 0\n3 4 2 1\nv 0 1\nv 1 2\nv 2 1\ne 0 1 1\ne 1 0 1\ne 1 2 1\ne 2 1 1\nt # -1",
                     GFSM_MODE_VLABEL) == GFSM_OK, "labelled");
    const char *bad[] = {
        "",                                                // empty file
        "t # 0\n",                                         // header missing
        "x # 0\n2 2 1 1\n",                                // wrong start
        "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 5 1\nt # -1",   // target id beyond V
        "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 0\nt # -1",   // edge label 0
        "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1 7 7 7 7 7\nt # -1",  // over-long line
        "t # 0\n4294967295 2 1 1\nv 0 1\nt # -1",           // absurd vertex count
        "t # 0\n2 99999999999 1 1\nv 0 1\nv 1 1\ne 0 1 1\nt # -1",  // absurd edge count
        "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1",      // truncated edge line, no end marker
        "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 99999999999999999999 1 1\nt # -1",  // overflowing number
        "t # 0\n\n\n\n",                                   // blank lines only
    };
    int k = 0;
    for (const char *t : bad) {
        const gfsm_status c = load_text(dir, "bad" + std::to_string(k++), t);
        expect(c == GFSM_ERR_INVALID_DATA || c == GFSM_ERR_MISSING_DATA || c == GFSM_ERR_ARG,
               "malformed text " + std::to_string(k - 1) + " -> " + std::to_string(c));
    }
    {
        HostGraph *g = nullptr;
        expect(load_tve((dir + "/missing.g").c_str(), GFSM_MODE_REF, g).code == GFSM_ERR_MISSING_FILE, "missing file");
        delete g;
    }

    // --- generated graphs: text round trip, image round trip, damaged images
    for (uint32_t mode : {GFSM_MODE_REF, GFSM_MODE_VLABEL}) {
        HostGraph *g = nullptr;
        expect(gen_rmat(12, 8, 0.57, 0.19, 0.19, 6, 1.0, 3, 7 + mode, mode, g).good(), "gen_rmat");
        std::unique_ptr<HostGraph> G(g);
        const std::string tp = dir + "/rmat" + std::to_string(mode) + ".g";
        expect(write_tve(*G, tp.c_str()).good(), "write_tve");
        HostGraph *g2 = nullptr;
        expect(load_tve(tp.c_str(), mode, g2).good(), "reload text");
        std::unique_ptr<HostGraph> G2(g2);
        expect(G2 && G2->nbr == G->nbr && G2->dir_ptr == G->dir_ptr, "text round trip keeps the layout");
        const std::string bp = dir + "/rmat" + std::to_string(mode) + ".bin";
        expect(save_bin(*G, bp.c_str()).good(), "save_bin");
        HostGraph *g3 = nullptr;
        expect(load_bin(bp.c_str(), g3).good(), "load_bin");
        std::unique_ptr<HostGraph> G3(g3);
        expect(G3 && G3->nbr == G->nbr && G3->rel_vtx == G->rel_vtx, "image round trip keeps the layout");
        const std::string img = read_file(bp);
        const size_t hdr = 8 + 4 * 4 + 6 * 8;
        std::vector<std::string> dmg;
        dmg.push_back(img.substr(0, img.size() - 5));                    // truncated
        dmg.push_back(img.substr(0, 20));                                // short header
        dmg.push_back(img + std::string(8, '\0'));                       // extended
        for (size_t at : {hdr + 3, hdr + 4097, img.size() / 2, img.size() - 1}) {
            std::string d = img;
            d[at] ^= 0x5a;                                               // payload damage
            dmg.push_back(d);
        }
        {
            std::string d = img;                                         // degree beyond 2^26
            const uint32_t big = 1u << 26;
            std::memcpy(&d[20], &big, 4);
            dmg.push_back(d);
        }
        {
            std::string d = img;                                         // counts that wrap a byte sum
            const uint64_t a = 1ull << 62, b = 0ull - (1ull << 62);
            std::memcpy(&d[24], &a, 8);
            std::memcpy(&d[40], &b, 8);
            dmg.push_back(d);
        }
        {
            std::string d = img;                                         // V that disagrees with the arrays
            const uint32_t v = 3;
            std::memcpy(&d[16], &v, 4);
            dmg.push_back(d);
        }
        int j = 0;
        for (const std::string &d : dmg) {
            const std::string p = dir + "/dmg" + std::to_string(mode) + "_" + std::to_string(j++) + ".bin";
            write_file(p, d);
            HostGraph *h = nullptr;
            const Status st = load_bin(p.c_str(), h);
            delete h;
            expect(st.code == GFSM_ERR_INVALID_DATA, "damaged image " + std::to_string(j - 1) + " -> " +
                                                         std::to_string(st.code));
        }
        // --- the planner on generated queries, both orders, and the projector
        for (uint32_t qs = 1; qs <= 12; qs++) {
            for (uint32_t kind = 0; kind < 2; kind++) {
                HostGraph *qg = nullptr;
                if (!gen_query_walk(*G, qs, kind, 100 * qs + kind, qg).good()) continue;
                std::unique_ptr<HostGraph> Q(qg);
                Plan p1, p2;
                expect(plan_query(*Q, p1).good(), "plan_query");
                expect(plan_query(*Q, p2, G.get()).good(), "plan_query by cost");
                expect(p1.q == Q->V && p2.q == Q->V, "plan covers every query vertex");
                if (p1.q >= 2) {
                    std::vector<double> nodes(p1.q), cand(p1.q);
                    project_levels(*G, p1, 100, 1000, nodes.data(), cand.data());
                    for (uint32_t d = 0; d < p1.q; d++) expect(nodes[d] >= 0 && cand[d] >= 0, "projection");
                }
            }
        }
    }
    // --- a disconnected query and a query with a vertex reached only by in-entries
    {
        const std::string p = dir + "/disc.g";
        write_file(p, "t # 0\n4 2 1 1\nv 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 1 0 1\nt # -1");
        HostGraph *g = nullptr;
        expect(load_tve(p.c_str(), GFSM_MODE_REF, g).good(), "load disconnected");
        std::unique_ptr<HostGraph> Q(g);
        Plan pl;
        expect(plan_query(*Q, pl).code == GFSM_ERR_QUERY, "disconnected query rejected");
    }
    std::fprintf(stderr, "host sanitizer run: %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
