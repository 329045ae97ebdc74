// graphio.cpp -- versioned binary image of a built graph (no reference
// counterpart: the reference re-parses the t/v/e text and rebuilds its CCSR
// on every run, HnHelper.cuh:283-499, which dominates end-to-end time on
// large graphs).
//
// Layout (little-endian):
//   header   magic "GFSMBIN1", u32 version, u32 mode, u32 V, u32 max_degree,
//            u64 ne (input entries), u64 nent, u64 ndir (directory entries,
//            sentinel excluded), u64 nrel (relations), u64 nrv (relation
//            index entries), u64 checksum of the payload
//   payload  vlabel[V], src[ne], dst[ne], elabel[ne]   (the input, t/v/e order)
//            nbr[nent], dir_ptr[V+1], dir[ndir+1] as {relation, begin},
//            rel_keys[nrel], rel_off[nrel+1] (u64), rel_vtx[nrv],
//            rel_beg[nrv], rel_cnt[nrv]
// Loading reads the arrays back and checks sizes, the checksum and the
// layout's invariants, so a truncated or corrupted image is rejected (233)
// rather than matched.
#include "host.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>

namespace gfsm {

namespace {

constexpr char kMagic[8] = {'G', 'F', 'S', 'M', 'B', 'I', 'N', '1'};
constexpr uint32_t kVersion = 1;

struct BinHeader {
    char magic[8];
    uint32_t version, mode, V, max_degree;
    uint64_t ne, nent, ndir, nrel, nrv, checksum;
};

// Order-dependent 64-bit checksum of byte ranges: each 1 MiB block is hashed
// on its own (SplitMix64 finaliser over its 8-byte words, the tail
// zero-padded), in parallel, and the block hashes are chained in order.
struct Checksum {
    static constexpr size_t kBlock = 1u << 20;
    uint64_t h = 0x243F6A8885A308D3ull;
    static uint64_t block_hash(const unsigned char *b, size_t bytes)
    {
        uint64_t x = 0x452821E638D01377ull;
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            x = mix64(x ^ w) + 0x9E3779B97F4A7C15ull;
        }
        if (i < bytes) {
            uint64_t w = 0;
            std::memcpy(&w, b + i, bytes - i);
            x = mix64(x ^ w) + 0x9E3779B97F4A7C15ull;
        }
        return x;
    }
    void add(const void *p, size_t bytes)
    {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        const size_t nb = (bytes + kBlock - 1) / kBlock;
        std::vector<uint64_t> part(nb);
        parallel_for(nb * 4096, [&](uint64_t lo, uint64_t hi) {  // scaled so small inputs still split
            for (uint64_t k = (lo + 4095) / 4096; k < (hi + 4095) / 4096; k++)
                part[k] = block_hash(b + k * kBlock, std::min(kBlock, bytes - k * kBlock));
        });
        for (uint64_t v : part) h = mix64(h ^ v) + 0x9E3779B97F4A7C15ull;
        h = mix64(h ^ bytes);
    }
};

template <class T>
void sum_vec(Checksum &c, const std::vector<T> &v)
{
    c.add(v.data(), v.size() * sizeof(T));
}

struct FileCloser {
    void operator()(FILE *f) const
    {
        if (f) std::fclose(f);
    }
};

template <class T>
bool write_vec(FILE *f, const std::vector<T> &v)
{
    return v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
}

template <class T>
bool read_vec(FILE *f, std::vector<T> &v, uint64_t n)
{
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

uint64_t payload_checksum(const HostGraph &g)
{
    Checksum c;
    sum_vec(c, g.vlabel);
    sum_vec(c, g.src);
    sum_vec(c, g.dst);
    sum_vec(c, g.elabel);
    sum_vec(c, g.nbr);
    sum_vec(c, g.dir_ptr);
    sum_vec(c, g.dir);
    sum_vec(c, g.rel_keys);
    sum_vec(c, g.rel_off);
    sum_vec(c, g.rel_vtx);
    sum_vec(c, g.rel_beg);
    sum_vec(c, g.rel_cnt);
    return c.h;
}

// The layout invariants a kernel relies on for in-bounds reads.
bool layout_ok(const HostGraph &g)
{
    const uint64_t ndir = g.dir.size() - 1, nent = g.nbr.size();
    if (g.V == 0 || nent != g.src.size() || nent >= 0xFFFFFFFFull) return false;
    if (g.dir_ptr.size() != (uint64_t)g.V + 1 || g.dir_ptr[0] != 0 || g.dir_ptr[g.V] != ndir) return false;
    if (g.dir[0].beg != 0 || g.dir[ndir].beg != nent) return false;
    std::atomic<bool> bad{false};
    std::atomic<uint32_t> maxdeg{0};
    for (uint32_t x = 0; x < g.V; x++)  // monotone first: the row checks index dir through it
        if (g.dir_ptr[x] > g.dir_ptr[x + 1]) return false;
    parallel_for(g.V, [&](uint64_t b, uint64_t e) {
        uint32_t md = 0;
        for (uint64_t x = b; x < e; x++) {
            for (uint32_t j = g.dir_ptr[x]; j + 1 < g.dir_ptr[x + 1]; j++)
                if (g.dir[j].rel >= g.dir[j + 1].rel) bad = true;  // relations ascend within a row
            md = std::max(md, g.dir[g.dir_ptr[x + 1]].beg - g.dir[g.dir_ptr[x]].beg);
        }
        uint32_t cur = maxdeg.load();
        while (md > cur && !maxdeg.compare_exchange_weak(cur, md)) {
        }
    });
    parallel_for(ndir, [&](uint64_t b, uint64_t e) {
        for (uint64_t j = b; j < e; j++)
            if (g.dir[j].beg >= g.dir[j + 1].beg) bad = true;  // every segment is non-empty
    });
    parallel_for(nent, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            if (g.nbr[i] >= g.V || g.src[i] >= g.V || g.dst[i] >= g.V) bad = true;
    });
    if (bad || maxdeg != g.max_degree || g.max_degree >= (1u << 26)) return false;  // build_graph's bound
    for (size_t k = 1; k < g.rel_keys.size(); k++)
        if (g.rel_keys[k - 1] >= g.rel_keys[k]) return false;
    if (g.rel_off.size() != g.rel_keys.size() + 1 || g.rel_off[0] != 0 || g.rel_off.back() != g.rel_vtx.size())
        return false;
    for (size_t k = 0; k + 1 < g.rel_off.size(); k++)
        if (g.rel_off[k] > g.rel_off[k + 1]) return false;
    parallel_for(g.rel_vtx.size(), [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++)
            if (g.rel_vtx[i] >= g.V || (uint64_t)g.rel_beg[i] + g.rel_cnt[i] > nent) bad = true;
    });
    return !bad;
}

}  // namespace

Status save_bin(const HostGraph &g, const char *path)
{
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "wb"));
    if (!f) return Status::err(GFSM_ERR_MISSING_FILE, std::string("cannot write ") + path);
    BinHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.mode = g.mode;
    h.V = g.V;
    h.max_degree = g.max_degree;
    h.ne = g.src.size();
    h.nent = g.nbr.size();
    h.ndir = g.dir.empty() ? 0 : g.dir.size() - 1;
    h.nrel = g.rel_keys.size();
    h.nrv = g.rel_vtx.size();
    h.checksum = payload_checksum(g);
    bool ok = std::fwrite(&h, sizeof(h), 1, f.get()) == 1;
    ok = ok && write_vec(f.get(), g.vlabel) && write_vec(f.get(), g.src) && write_vec(f.get(), g.dst) &&
         write_vec(f.get(), g.elabel) && write_vec(f.get(), g.nbr) && write_vec(f.get(), g.dir_ptr) &&
         write_vec(f.get(), g.dir) && write_vec(f.get(), g.rel_keys) && write_vec(f.get(), g.rel_off) &&
         write_vec(f.get(), g.rel_vtx) && write_vec(f.get(), g.rel_beg) && write_vec(f.get(), g.rel_cnt);
    if (!ok || std::fflush(f.get()) != 0) return Status::err(GFSM_ERR_MISSING_FILE, std::string("write failed: ") + path);
    return Status::ok();
}

Status load_bin(const char *path, HostGraph *&out)
{
    out = nullptr;
    std::unique_ptr<FILE, FileCloser> f(std::fopen(path, "rb"));
    if (!f) return Status::err(GFSM_ERR_MISSING_FILE, std::string("cannot open ") + path);
    BinHeader h{};
    if (std::fread(&h, sizeof(h), 1, f.get()) != 1 || std::memcmp(h.magic, kMagic, 8) != 0)
        return Status::err(GFSM_ERR_INVALID_DATA, "not a graph image");
    if (h.version != kVersion) return Status::err(GFSM_ERR_INVALID_DATA, "unsupported graph image version");
    if (h.mode != GFSM_MODE_REF && h.mode != GFSM_MODE_VLABEL)
        return Status::err(GFSM_ERR_INVALID_DATA, "bad mode in graph image");
    // Every count fits 32 bits (the layout's indices are u32) and the degree
    // bound of build_graph holds, so the size sum below cannot wrap.
    if (h.ne >= (1ull << 32) || h.nent >= 0xFFFFFFFFull || h.ndir >= (1ull << 32) || h.nrel >= (1ull << 32) ||
        h.nrv >= (1ull << 32) || h.max_degree >= (1u << 26))
        return Status::err(GFSM_ERR_INVALID_DATA, "graph image header out of range");
    // Sizes must add up to the file length before anything large is allocated.
    const uint64_t want = sizeof(h) + 4ull * h.V + 12ull * h.ne + 4ull * h.nent + 4ull * ((uint64_t)h.V + 1) +
                          8ull * (h.ndir + 1) + 4ull * h.nrel + 8ull * (h.nrel + 1) + 12ull * h.nrv;
    if (std::fseek(f.get(), 0, SEEK_END) != 0) return Status::err(GFSM_ERR_INVALID_DATA, "cannot size graph image");
    const long long len = std::ftell(f.get());
    if (len < 0 || (uint64_t)len != want) return Status::err(GFSM_ERR_INVALID_DATA, "graph image size mismatch");
    if (std::fseek(f.get(), (long)sizeof(h), SEEK_SET) != 0) return Status::err(GFSM_ERR_INVALID_DATA, "cannot read graph image");
    std::unique_ptr<HostGraph> g(new HostGraph);
    g->mode = h.mode;
    g->V = h.V;
    g->max_degree = h.max_degree;
    bool ok = read_vec(f.get(), g->vlabel, h.V) && read_vec(f.get(), g->src, h.ne) &&
              read_vec(f.get(), g->dst, h.ne) && read_vec(f.get(), g->elabel, h.ne) &&
              read_vec(f.get(), g->nbr, h.nent) && read_vec(f.get(), g->dir_ptr, (uint64_t)h.V + 1) &&
              read_vec(f.get(), g->dir, h.ndir + 1) && read_vec(f.get(), g->rel_keys, h.nrel) &&
              read_vec(f.get(), g->rel_off, h.nrel + 1) && read_vec(f.get(), g->rel_vtx, h.nrv) &&
              read_vec(f.get(), g->rel_beg, h.nrv) && read_vec(f.get(), g->rel_cnt, h.nrv);
    if (!ok) return Status::err(GFSM_ERR_INVALID_DATA, "truncated graph image");
    if (payload_checksum(*g) != h.checksum) return Status::err(GFSM_ERR_INVALID_DATA, "graph image checksum mismatch");
    if (!layout_ok(*g)) return Status::err(GFSM_ERR_INVALID_DATA, "graph image layout is inconsistent");
    out = g.release();
    return Status::ok();
}

}  // namespace gfsm
