// Sorted mer list files (.sslist, format version 5): the on-disk layout of
// FileSML (LM/FileSML.cpp:401-459 writes it, LM/FileSML.cpp:120-192 reads it
// back with LoadFile2's status codes).
//
// File layout (little-endian, x86-64 struct layout of SMLHeader,
// LM/SortedMerList.h:48-63):
//   offset    0  uint32 version            5 (DNAFileSML::FormatVersion)
//             4  uint32 alphabet_bits      2
//             8  uint64 seed
//            16  uint32 seed_length
//            20  uint32 seed_weight
//            24  uint64 length             sequence length
//            32  uint32 unique_mers        0xFFFFFFFF (not counted)
//            36  uint32 word_size          32 (bits per packed word)
//            40  uint8  little_endian      1
//            42  int16  id                 0
//            44  uint8  circular           0
//            45  uint8  translation_table[255]   BasicDNATable
//           300  char   description[2048]  empty
//          2352  uint32 sequence[ceil(2*length/32) + 2]   2-bit packing, MSB first
//          then  uint32 positions[SMLLength]              the sorted list
// Header defaults as SortedMerList's constructor and Create set them
// (LM/SortedMerList.cpp:95-105, 786-824).  The reference leaves word_size,
// little_endian, the description after its first byte and the two padding
// words of the sequence unwritten; this writer fills them with the constants
// above and zeros.  Readers of either ignore them.
#include "mh_internal.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

using namespace mh;

namespace {

constexpr uint64_t HEADER_BYTES = 2352;
constexpr uint32_t FORMAT_VERSION = 5;   // LM/DNAFileSML.h:58-62
constexpr uint32_t NO_UNIQUE_COUNT = 0xFFFFFFFFu;

void put32(uint8_t *b, uint64_t off, uint32_t v) { std::memcpy(b + off, &v, 4); }
void put64(uint8_t *b, uint64_t off, uint64_t v) { std::memcpy(b + off, &v, 8); }
uint32_t get32(const uint8_t *b, uint64_t off) { uint32_t v; std::memcpy(&v, b + off, 4); return v; }
uint64_t get64(const uint8_t *b, uint64_t off) { uint64_t v; std::memcpy(&v, b + off, 8); return v; }

// BasicDNATable, LM/SortedMerList.cpp:29-47: C/c/B/b/Y/y -> 1,
// G/g/S/s/K/k -> 2, T/t -> 3, every other byte (A included) -> 0.
void basic_dna_table(uint8_t *t)
{
    std::memset(t, 0, 255);
    for (const char *p = "cCbByY"; *p; ++p) t[(uint8_t)*p] = 1;
    for (const char *p = "gGsSkK"; *p; ++p) t[(uint8_t)*p] = 2;
    for (const char *p = "tT"; *p; ++p) t[(uint8_t)*p] = 3;
}

struct File {
    std::FILE *f = nullptr;
    explicit File(std::FILE *p) : f(p) {}
    ~File() { if (f) std::fclose(f); }
};

int load_status(int *status, int code, int rc, const std::string &msg)
{
    if (status) *status = code;
    set_error(msg);
    return rc;
}

}  // namespace

extern "C" {

int mh_sml_save(mh_context *ctx, const mh_sml *s, const char *path)
{
    MH_CHECK(ctx && s && path, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    uint8_t hdr[HEADER_BYTES];
    std::memset(hdr, 0, sizeof hdr);
    put32(hdr, 0, FORMAT_VERSION);
    put32(hdr, 4, 2);
    put64(hdr, 8, s->seed);
    put32(hdr, 16, (uint32_t)s->seed_length);
    put32(hdr, 20, (uint32_t)s->seed_weight);
    put64(hdr, 24, s->len);
    put32(hdr, 32, NO_UNIQUE_COUNT);
    put32(hdr, 36, 32);
    hdr[40] = 1;
    basic_dna_table(hdr + 45);
    std::vector<uint32_t> words(s->n_words), pos(s->n);
    if (s->n_words)
        MH_HIP(hipMemcpyAsync(words.data(), s->d_words, s->n_words * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (s->n) MH_HIP(hipMemcpyAsync(pos.data(), s->d_pos, s->n * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    File out(std::fopen(path, "wb"));
    MH_CHECK(out.f, MH_ERR_IO, std::string("cannot open for writing: ") + path);
    bool ok = std::fwrite(hdr, 1, HEADER_BYTES, out.f) == HEADER_BYTES;
    ok = ok && std::fwrite(words.data(), 4, words.size(), out.f) == words.size();
    ok = ok && std::fwrite(pos.data(), 4, pos.size(), out.f) == pos.size();
    ok = std::fclose(out.f) == 0 && ok;
    out.f = nullptr;
    MH_CHECK(ok, MH_ERR_IO, std::string("Error writing sorted mer list to disk: ") + path);
    return MH_OK;
}

int mh_sml_load(mh_context *ctx, const char *path, mh_sml **out, int *status)
{
    MH_CHECK(ctx && path && out, MH_ERR_INVALID, "null argument");
    if (status) *status = 0;
    *out = nullptr;
    MH_HIP(hipSetDevice(ctx->device));
    File in(std::fopen(path, "rb"));
    if (!in.f) return load_status(status, 1, MH_ERR_IO, std::string("Unable to open file: ") + path);
    uint8_t hdr[HEADER_BYTES];
    if (std::fread(hdr, 1, HEADER_BYTES, in.f) != HEADER_BYTES)
        return load_status(status, 2, MH_ERR_IO, "Unable to read file (short header)");
    if (get32(hdr, 0) != FORMAT_VERSION) return load_status(status, 3, MH_ERR_IO, "Unsupported file format");
    const uint32_t abits = get32(hdr, 4);
    const uint64_t seed = get64(hdr, 8);
    const uint32_t sl = get32(hdr, 16), sw = get32(hdr, 20);
    const uint64_t len = get64(hdr, 24);
    const uint8_t circular = hdr[44];
    // what the reference would accept but this build does not represent
    if (abits != 2 || circular)
        return load_status(status, 6, MH_ERR_UNSUPPORTED, "only 2-bit DNA lists of linear sequences are supported");
    if (seed == 0 || (int)sl != seed_length((int64_t)seed) || (int)sw != seed_weight((int64_t)seed) || sl > 31)
        return load_status(status, 6, MH_ERR_UNSUPPORTED, "seed fields of the header are inconsistent");
    if (len >= (1ull << 32))
        return load_status(status, 6, MH_ERR_UNSUPPORTED, "sequence longer than 2^32 (32-bit SML positions)");
    const uint64_t n_words = (len * 2 + 31) / 32 + 2;
    std::vector<uint32_t> words(n_words);
    if (std::fread(words.data(), 4, n_words, in.f) != n_words)
        return load_status(status, 4, MH_ERR_IO, "Error reading sequence data");
    // bits past the sequence end are not sequence (the reference leaves the
    // padding words unwritten): clear them
    {
        const uint64_t bits = len * 2;
        const uint64_t w = bits / 32, r = bits % 32;
        if (r) words[w] &= ~0u << (32 - r);
        for (uint64_t q = w + (r ? 1 : 0); q < n_words; ++q) words[q] = 0;
    }
    // SMLLength, LM/SortedMerList.cpp:288-295 (linear sequence).  The
    // reference's own check seeks 8*length bytes past the sequence, which
    // succeeds past the end of the file; here the positions must all be there.
    const uint64_t n = len >= sl ? len - sl + 1 : 0;
    std::vector<uint32_t> pos(n);
    if (std::fread(pos.data(), 4, n, in.f) != n) return load_status(status, 5, MH_ERR_IO, "Premature end of file");
    // the device kernels index the sequence by position: each must be a
    // distinct position below SMLLength
    {
        std::vector<uint8_t> seen((n + 7) / 8, 0);
        for (uint64_t x = 0; x < n; ++x) {
            const uint32_t p = pos[x];
            if (p >= n || (seen[p >> 3] >> (p & 7)) & 1)
                return load_status(status, 7, MH_ERR_INVALID, "positions are not a permutation of [0, SMLLength)");
            seen[p >> 3] |= (uint8_t)(1u << (p & 7));
        }
    }
    mh_sml *s = new mh_sml();
    s->ctx = ctx;
    ctx->smls.insert(s);
    s->len = len;
    s->n = n;
    s->seed = seed;
    s->seed_length = (int)sl;
    s->seed_weight = (int)sw;
    s->n_words = n_words;
    int unsorted = 0;
    int rc = sml_alloc(ctx, s);
    if (rc == MH_OK) {
        hipError_t e = hipMemcpyAsync(s->d_words, words.data(), n_words * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess && n) e = hipMemcpyAsync(s->d_pos, pos.data(), n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
            rc = MH_ERR_DEVICE;
        }
    }
    if (rc == MH_OK) rc = sml_rebuild_keys(ctx, s, &unsorted);
    if (rc == MH_OK && unsorted) rc = load_status(status, 7, MH_ERR_INVALID, "list is not sorted by seed key");
    if (rc != MH_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        mh_sml_free(s);
        return rc;
    }
    *out = s;
    return MH_OK;
}

}  // extern "C"
