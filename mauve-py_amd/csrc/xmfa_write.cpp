// XMFA output: WriteStandardAlignment, LM/IntervalList.h:353-442.
//
//   #FormatVersion Mauve1
//   #Sequence<i>File\t<name>          for each genome i (1-based)
//   #Sequence<i>Entry\t<i>            only when every genome names the same file
//   #Sequence<i>Format\t<format>      from the name's extension (below)
//   [#Annotation<i>File / Format      for DNAstar and GenBank names]
//   [#BackboneFile\t<name>]
//   then per interval of non-zero alignment length, per genome:
//   > <i>:<left>-<right> <+|-> <name>
//   <aligned row in lines of 80>
//   and "=" after each interval.
// A genome absent from an interval (start 0) is written as "0-0 +" in the
// first interval and left out of later ones.  The format follows the source
// class the reference's gnSourceFactory picks by the text after the last '.'
// (libGenome/gnSourceFactory.cpp:34-46, 88-95): .fas/.FAS FastA, .seq/.SEQ
// DNAstar, .gbk/.GBK/.gb/.GB GenBank, .raw/.RAW raw, anything else FastA.
#include "mh_internal.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>

using namespace mh;

namespace {

enum SourceKind { FASTA, RAW, DNASTAR, GENBANK };

SourceKind source_kind(const std::string &name)
{
    const size_t dot = name.rfind('.');
    if (dot == std::string::npos) return FASTA;
    const std::string ext = name.substr(dot);
    if (ext == ".raw" || ext == ".RAW") return RAW;
    if (ext == ".seq" || ext == ".SEQ") return DNASTAR;
    if (ext == ".gbk" || ext == ".GBK" || ext == ".gb" || ext == ".GB") return GENBANK;
    return FASTA;
}

}  // namespace

extern "C" {

int mh_xmfa_write(const char *path, int seq_count, const char *const *seq_files, const char *backbone_file,
                  const mh_xmfa_row *rows, uint64_t n_intervals)
{
    MH_CHECK(path && seq_count >= 1 && (rows || n_intervals == 0), MH_ERR_INVALID, "null argument");
    std::FILE *fp = std::fopen(path, "wb");
    MH_CHECK(fp, MH_ERR_IO, std::string("Error opening \"") + path + "\"");
    std::unique_ptr<std::FILE, int (*)(std::FILE *)> closer(fp, std::fclose);
    if (n_intervals == 0) return MH_OK;
    std::vector<std::string> names(seq_count);
    for (int s = 0; s < seq_count; ++s) names[s] = seq_files && seq_files[s] ? seq_files[s] : "";
    bool single = true;
    for (int s = 1; s < seq_count; ++s)
        if (names[s] != names[0]) {
            single = false;
            break;
        }
    std::ostringstream head;
    head << "#FormatVersion Mauve1\n";
    for (int s = 0; s < seq_count; ++s) {
        head << "#Sequence" << s + 1 << "File\t" << names[s] << "\n";
        if (single) head << "#Sequence" << s + 1 << "Entry\t" << s + 1 << "\n";
        switch (source_kind(names[s])) {
        case FASTA: head << "#Sequence" << s + 1 << "Format\tFastA\n"; break;
        case RAW: head << "#Sequence" << s + 1 << "Format\traw\n"; break;
        case DNASTAR:
            head << "#Sequence" << s + 1 << "Format\tDNAstar\n";
            head << "#Annotation" << s + 1 << "File\t" << names[s] << "\n";
            head << "#Annotation" << s + 1 << "Format\tDNAstar\n";
            break;
        case GENBANK:
            head << "#Sequence" << s + 1 << "Format\tGenBank\n";
            head << "#Annotation" << s + 1 << "File\t" << names[s] << "\n";
            head << "#Annotation" << s + 1 << "Format\tGenBank\n";
            break;
        }
    }
    if (backbone_file && backbone_file[0]) head << "#BackboneFile\t" << backbone_file << "\n";
    // the file as text pieces (headers) and rows (80 columns a line); rows
    // are laid out into one buffer in parallel, then the buffer is written
    struct Piece {
        std::string text;
        const char *row = nullptr;
        uint64_t row_len = 0, at = 0;
    };
    std::vector<Piece> pieces;
    pieces.push_back(Piece{head.str()});
    for (uint64_t iv = 0; iv < n_intervals; ++iv) {
        const mh_xmfa_row *r = rows + iv * (uint64_t)seq_count;
        // an interval's alignment length is its rows' length (equal for all)
        if (r[0].row_len == 0) continue;
        for (int s = 0; s < seq_count; ++s) {
            MH_CHECK(r[s].row_len == r[0].row_len && (r[s].row || r[s].row_len == 0), MH_ERR_INVALID,
                     "rows of an interval differ in length");
            const int64_t st = r[s].start;
            if (st == 0 && iv > 0) continue;
            std::ostringstream h;
            h << "> " << s + 1 << ":";
            const uint64_t a = st < 0 ? (uint64_t)(-st) : (uint64_t)st;
            if (st > 0)
                h << a << "-" << a + r[s].length - 1 << " + ";
            else if (st == 0)
                h << 0 << "-" << 0 << " + ";
            else
                h << a << "-" << a + r[s].length - 1 << " - ";
            h << (single ? names[0] : names[s]) << "\n";
            pieces.push_back(Piece{h.str()});
            Piece rp;
            rp.row = r[s].row;
            rp.row_len = r[s].row_len;
            pieces.push_back(rp);
        }
        pieces.push_back(Piece{"=\n"});
    }
    uint64_t total = 0;
    for (Piece &pc : pieces) {
        pc.at = total;
        total += pc.row ? pc.row_len + (pc.row_len + 79) / 80 : pc.text.size();
    }
    static std::mutex mu;
    static HostBuf buf;
    std::lock_guard<std::mutex> lock(mu);
    char *out = buf.get<char>(total);
    MH_CHECK(out, MH_ERR_NOMEM, "out of host memory");
    // tasks: header pieces one each, rows in blocks of 8192 lines
    constexpr uint64_t LINES = 8192;
    std::vector<std::pair<size_t, uint64_t>> tasks;
    for (size_t k = 0; k < pieces.size(); ++k) {
        if (!pieces[k].row) {
            std::memcpy(out + pieces[k].at, pieces[k].text.data(), pieces[k].text.size());
            continue;
        }
        for (uint64_t l = 0; l * 80 < pieces[k].row_len; l += LINES) tasks.push_back({k, l});
    }
    host_parallel(tasks.size(), [&](size_t t) {
        const Piece &pc = pieces[tasks[t].first];
        const uint64_t l0 = tasks[t].second, l1 = std::min((pc.row_len + 79) / 80, l0 + LINES);
        char *w = out + pc.at + l0 * 81;
        for (uint64_t l = l0; l < l1; ++l) {
            const uint64_t p = l * 80, n = std::min<uint64_t>(80, pc.row_len - p);
            std::memcpy(w, pc.row + p, n);
            w[n] = '\n';
            w += n + 1;
        }
    });
    MH_CHECK(std::fwrite(out, 1, total, fp) == total, MH_ERR_IO, std::string("Error writing alignment: ") + path);
    MH_CHECK(std::fclose(closer.release()) == 0, MH_ERR_IO, std::string("Error writing alignment: ") + path);
    return MH_OK;
}

}  // extern "C"
