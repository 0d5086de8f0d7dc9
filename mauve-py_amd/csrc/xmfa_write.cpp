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

#include <cstring>
#include <fstream>

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
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    MH_CHECK(out.is_open(), MH_ERR_IO, std::string("Error opening \"") + path + "\"");
    if (n_intervals == 0) return MH_OK;
    std::vector<std::string> names(seq_count);
    for (int s = 0; s < seq_count; ++s) names[s] = seq_files && seq_files[s] ? seq_files[s] : "";
    bool single = true;
    for (int s = 1; s < seq_count; ++s)
        if (names[s] != names[0]) {
            single = false;
            break;
        }
    out << "#FormatVersion Mauve1\n";
    for (int s = 0; s < seq_count; ++s) {
        out << "#Sequence" << s + 1 << "File\t" << names[s] << "\n";
        if (single) out << "#Sequence" << s + 1 << "Entry\t" << s + 1 << "\n";
        switch (source_kind(names[s])) {
        case FASTA: out << "#Sequence" << s + 1 << "Format\tFastA\n"; break;
        case RAW: out << "#Sequence" << s + 1 << "Format\traw\n"; break;
        case DNASTAR:
            out << "#Sequence" << s + 1 << "Format\tDNAstar\n";
            out << "#Annotation" << s + 1 << "File\t" << names[s] << "\n";
            out << "#Annotation" << s + 1 << "Format\tDNAstar\n";
            break;
        case GENBANK:
            out << "#Sequence" << s + 1 << "Format\tGenBank\n";
            out << "#Annotation" << s + 1 << "File\t" << names[s] << "\n";
            out << "#Annotation" << s + 1 << "Format\tGenBank\n";
            break;
        }
    }
    if (backbone_file && backbone_file[0]) out << "#BackboneFile\t" << backbone_file << "\n";
    for (uint64_t iv = 0; iv < n_intervals; ++iv) {
        const mh_xmfa_row *r = rows + iv * (uint64_t)seq_count;
        // an interval's alignment length is its rows' length (equal for all)
        if (r[0].row_len == 0) continue;
        for (int s = 0; s < seq_count; ++s) {
            MH_CHECK(r[s].row_len == r[0].row_len && (r[s].row || r[s].row_len == 0), MH_ERR_INVALID,
                     "rows of an interval differ in length");
            const int64_t st = r[s].start;
            if (st == 0 && iv > 0) continue;
            out << "> " << s + 1 << ":";
            const uint64_t a = st < 0 ? (uint64_t)(-st) : (uint64_t)st;
            if (st > 0)
                out << a << "-" << a + r[s].length - 1 << " + ";
            else if (st == 0)
                out << 0 << "-" << 0 << " + ";
            else
                out << a << "-" << a + r[s].length - 1 << " - ";
            out << (single ? names[0] : names[s]) << "\n";
            for (uint64_t p = 0; p < r[s].row_len; p += 80) {
                const uint64_t w = p + 80 < r[s].row_len ? 80 : r[s].row_len - p;
                out.write(r[s].row + p, (std::streamsize)w);
                out << "\n";
            }
        }
        out << "=\n";
    }
    out.flush();
    MH_CHECK(out.good(), MH_ERR_IO, std::string("Error writing alignment: ") + path);
    return MH_OK;
}

}  // extern "C"
