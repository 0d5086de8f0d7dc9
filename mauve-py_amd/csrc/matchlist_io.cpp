// Match-list files (format version 3), the text `progressiveMauve --mums`
// writes and `--match-input` reads: WriteList and ReadList,
// LM/MatchList.h:524-662.
//
//   FormatVersion\t3
//   SequenceCount\t<k>
//   Sequence<i>File\t<name, or null>        } for i in 0..k-1
//   Sequence<i>Length\t<length, or 0>       }
//   MatchCount\t<n>
//   <length>\t<start 0>\t...\t<start k-1>\t<id>\t0\t0     one line per match
//
// Starts are 1-based, negative on the reverse strand, 0 where a genome is
// absent.  <id> is the match's address in the reference (a void* printed in
// hexadecimal); this writer prints 0x<index + 1>, which the reference's
// reader accepts as well (it only keys a map with it).  An empty list writes
// nothing, as WriteList does.
#include "mh_internal.hpp"

#include <fstream>
#include <sstream>

using namespace mh;

extern "C" {

int mh_match_list_write(const char *path, const mh_match *m, uint64_t n, int seq_count,
                        const char *const *seq_files, const uint64_t *seq_lengths)
{
    MH_CHECK(path && (m || n == 0), MH_ERR_INVALID, "null argument");
    MH_CHECK(seq_count >= 1 && seq_count <= MH_MAX_SEQ, MH_ERR_INVALID, "sequence count out of range");
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    MH_CHECK(out.is_open(), MH_ERR_IO, std::string("Error opening \"") + path + "\"");
    if (n) {
        out << "FormatVersion" << '\t' << 3 << "\n";
        out << "SequenceCount" << '\t' << seq_count << "\n";
        for (int q = 0; q < seq_count; ++q) {
            out << "Sequence" << q << "File" << '\t';
            if (seq_files && seq_files[q])
                out << seq_files[q];
            else
                out << "null";
            out << "\n";
            out << "Sequence" << q << "Length" << '\t';
            if (seq_lengths)
                out << seq_lengths[q];
            else
                out << "0";
            out << "\n";
        }
        out << "MatchCount" << '\t' << n << "\n";
        for (uint64_t x = 0; x < n; ++x) {
            out << m[x].length;
            for (int q = 0; q < seq_count; ++q) out << '\t' << m[x].start[q];
            out << '\t' << "0x" << std::hex << (x + 1) << std::dec << '\t' << 0 << '\t' << 0 << "\n";
        }
    }
    out.flush();
    MH_CHECK(out.good(), MH_ERR_IO, std::string("Error writing match list: ") + path);
    return MH_OK;
}

int mh_match_list_read(const char *path, mh_match_list **out)
{
    MH_CHECK(path && out, MH_ERR_INVALID, "null argument");
    *out = nullptr;
    std::ifstream in(path, std::ios::binary);
    MH_CHECK(in.is_open(), MH_ERR_IO, std::string("Error opening ") + path);
    const char *bad = "Possibly corrupt file or invalid file format";
    std::string tag;
    in >> tag;
    MH_CHECK(tag == "FormatVersion", MH_ERR_IO, bad);
    in >> tag;
    MH_CHECK(tag == "3", MH_ERR_IO, bad);
    in >> tag;
    MH_CHECK(tag == "SequenceCount", MH_ERR_IO, bad);
    unsigned seq_count = 0;
    in >> seq_count;
    MH_CHECK(in && seq_count >= 2, MH_ERR_IO, bad);
    MH_CHECK(seq_count <= MH_MAX_SEQ, MH_ERR_UNSUPPORTED, "match lists of more than 4 genomes");
    mh_match_list *l = new mh_match_list();
    l->seq_count = (int)seq_count;
    for (unsigned q = 0; q < seq_count; ++q) {
        in >> tag;                          // Sequence<q>File
        std::getline(in, tag);
        l->files.push_back(tag.empty() ? tag : tag.substr(1));   // drop the tab
        in >> tag;                          // Sequence<q>Length
        uint64_t len = 0;
        in >> len;
        l->lengths.push_back(len);
    }
    uint64_t count = 0;
    in >> tag >> count;
    std::string line;
    std::getline(in, line);
    bool ok = (bool)in && tag == "MatchCount";
    while (ok && std::getline(in, line)) {
        std::istringstream ls(line);
        mh_match mm{};
        ls >> mm.length;
        for (unsigned q = 0; q < seq_count; ++q) ls >> mm.start[q];
        void *id = nullptr;
        unsigned sub = 0, sup = 0;
        ls >> id >> sub >> sup;
        // subset and superset lists are not part of this format version
        ok = (bool)ls && sub == 0 && sup == 0;
        if (ok) l->v.push_back(mm);
    }
    if (!ok || l->v.size() != count) {
        delete l;
        set_error(bad);
        return MH_ERR_IO;
    }
    *out = l;
    return MH_OK;
}

int mh_match_list_seq_count(const mh_match_list *l) { return l ? l->seq_count : 0; }

const char *mh_match_list_seq_file(const mh_match_list *l, int i)
{
    return l && i >= 0 && i < (int)l->files.size() ? l->files[i].c_str() : nullptr;
}

uint64_t mh_match_list_seq_length(const mh_match_list *l, int i)
{
    return l && i >= 0 && i < (int)l->lengths.size() ? l->lengths[i] : 0;
}

}  // extern "C"
