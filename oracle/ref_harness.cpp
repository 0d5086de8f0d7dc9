// ref_harness.cpp -- driver around the reference's own SortedMerList and
// libMUSCLE code, compiled from the sources under the reference tree by
// oracle/Makefile into oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it is used
// to write golden vectors and to cross-check the plain-C oracle.  Nothing in
// the product library links it.
//
//   ref_harness keys <seed> <in.seq> <out.bin>
//       Runs DNAMemorySML::Create (LM/MemorySML.cpp:45-60) on the sequence
//       and writes: n_words, packed words (LM/SortedMerList.cpp:306-317),
//       n, fill keys in position order (FillDnaSML or FillDnaSeedSML),
//       GetDnaSeedMer per position (LM/SortedMerList.cpp:764-769), and the
//       sorted positions with their keys as the reference std::sort left them.
//
//   ref_harness findmer <seed> <in.seq> <queries.bin> <out.bin>
//       Runs DNAMemorySML::Create on the sequence, then SortedMerList::FindMer
//       (LM/SortedMerList.cpp:170-179, bsearch 380-394) for each uint64 query
//       of queries.bin, and writes per query the found flag (uint64) and the
//       index it left in result (uint64; UINT64_MAX if it left it unset).
//
//   ref_harness app <in.txt> <out.txt>
//       For each line "row1 row2" (two gapped rows of one length) sets up
//       MSAs as MuscleInterface::ProfileAlignFast(anchored = true) does
//       (LM/MuscleInterface.cpp:1053-1106) and writes three lines:
//         "S <n> <score bits ...>"  LetterObjScoreXP's column scores after
//                                   PrepareMSAforScoring (MU/anchoredpp.cpp:
//                                   256-316, 430-441), float bits in hex;
//         "A <k> <cols ...>"        FindAnchorColsPP (MU/anchoredpp.cpp:354-404);
//         "R <row1> <row2>"         AnchoredProfileProfile's output rows
//                                   (MU/anchoredpp.cpp:443-552).
//
//   ref_harness refinew <window> <in.txt> <out.txt>
//       For each line "row1 row2" runs MuscleInterface::RefineFast(ga,
//       window) for two sequences (LM/MuscleInterface.cpp:823-894):
//       the globals, the guide tree, RefineW (MU/refinew.cpp:75-170), and
//       writes the two output rows.
//
//   ref_harness nw <in.txt> <out.txt>
//       For each line "A B" runs ProfileProfile (MU/profile.cpp:68-93) on two
//       one-sequence MSAs with the globals set as MuscleInterface::
//       ProfileAlignFast sets them (LM/MuscleInterface.cpp:1090-1106), and
//       writes the two aligned rows.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <fstream>
#include <sstream>
#include <iostream>

#include "libGenome/gnSequence.h"
#include "libMems/DNAMemorySML.h"

#include "libMUSCLE/muscle.h"
#include "libMUSCLE/msa.h"
#include "libMUSCLE/profile.h"
#include "libMUSCLE/params.h"
#include "libMUSCLE/tree.h"

using namespace std;

// libMUSCLE functions with external linkage that no header declares
namespace muscle {
SCORE LetterObjScoreXP(const MSA &msa1, const MSA &msa2, SCORE MatchScore[]);
void FindAnchorColsPP(const MSA &msa1, const MSA &msa2, unsigned AnchorCols[], unsigned *ptruAnchorColCount);
void PrepareMSAforScoring(MSA &msa);
void RefineW(const MSA &msaIn, MSA &msaOut);
}

namespace {

// Exposes the protected state of the reference SML class to the driver.
class ProbeSML : public mems::DNAMemorySML {
public:
    const uint32* words() const { return sequence; }
    gnSeqI nwords() const { return binary_seq_len; }
    void fill(const genome::gnSequence& seq, vector<mems::bmer>& arr) {
        if (header.seed_length != header.seed_weight) FillDnaSeedSML(seq, arr);
        else FillSML(seq, arr);
    }
    uint64 seedmer(gnSeqI p) const { return GetDnaSeedMer(p); }
    gnSeqI npos() const { return positions.size(); }
    gnSeqI posat(gnSeqI i) const { return positions[i]; }
};

void put64(FILE* f, uint64_t v) { fwrite(&v, 8, 1, f); }

int do_keys(uint64_t seed, const char* in, const char* out)
{
    ifstream is(in);
    string s((istreambuf_iterator<char>(is)), istreambuf_iterator<char>());
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    genome::gnSequence seq(s);
    ProbeSML sml;
    sml.Create(seq, seed);
    FILE* f = fopen(out, "wb");
    if (!f) return 2;
    uint64_t nw = sml.nwords();
    // the reference leaves the padding words past the data uninitialised; the
    // driver reports only the words that hold characters
    uint64_t used = (s.size() * 2 + 31) / 32;
    put64(f, used);
    for (uint64_t i = 0; i < used; ++i) { uint32_t w = sml.words()[i]; fwrite(&w, 4, 1, f); }
    (void)nw;
    vector<mems::bmer> arr;
    sml.fill(seq, arr);
    uint64_t n = arr.size();
    put64(f, n);
    for (uint64_t i = 0; i < n; ++i) put64(f, arr[i].mer);
    for (uint64_t i = 0; i < n; ++i) put64(f, sml.seedmer(i));
    uint64_t np = sml.npos();
    put64(f, np);
    for (uint64_t i = 0; i < np; ++i) {
        uint32_t p = (uint32_t)sml.posat(i);
        fwrite(&p, 4, 1, f);
    }
    fclose(f);
    return 0;
}

int do_findmer(uint64_t seed, const char* in, const char* qpath, const char* out)
{
    ifstream is(in);
    string s((istreambuf_iterator<char>(is)), istreambuf_iterator<char>());
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    genome::gnSequence seq(s);
    ProbeSML sml;
    sml.Create(seq, seed);
    FILE* q = fopen(qpath, "rb");
    FILE* f = fopen(out, "wb");
    if (!q || !f) return 2;
    uint64_t v;
    while (fread(&v, 8, 1, q) == 1) {
        gnSeqI r = ~(gnSeqI)0;
        const bool found = sml.FindMer(v, r);
        put64(f, found ? 1 : 0);
        put64(f, (uint64_t)r);
    }
    fclose(q);
    fclose(f);
    return 0;
}

void msa_from_string(muscle::MSA& msa, const string& s, unsigned id_base)
{
    // same steps as msaFromSeqTable (LM/MuscleInterface.cpp:808-820)
    msa.SetSize(1, s.size());
    char name[32];
    snprintf(name, sizeof(name), "seq%u", 0u);
    msa.SetSeqName(0, name);
    msa.SetSeqId(0, id_base);
    for (size_t i = 0; i < s.size(); ++i) msa.SetChar(0, i, s[i]);
}

void set_pairwise_globals()
{
    using namespace muscle;
    // as ProfileAlignFast / RefineFast set them (LM/MuscleInterface.cpp:837-846, 1088-1097)
    g_SeqType.get() = SEQTYPE_DNA;
    g_uMaxIters.get() = 1;
    g_bStable.get() = true;
    g_bQuiet.get() = true;
    g_SeqWeight1.get() = SEQWEIGHT_ClustalW;
    SetMaxIters(g_uMaxIters.get());
    SetSeqWeightMethod(g_SeqWeight1.get());
}

// output row of msaOut for input sequence id `id` (GetSeqIndex order)
string row_of(const muscle::MSA& msa, unsigned id)
{
    for (unsigned r = 0; r < msa.GetSeqCount(); ++r)
        if (msa.GetSeqIndex(r) == id) {
            const char* buf = msa.GetSeqBuffer(msa.GetSeqIndex(r));
            return string(buf, msa.GetColCount());
        }
    return string();
}

int do_app(const char* in, const char* out)
{
    using namespace muscle;
    ifstream is(in);
    FILE* f = fopen(out, "w");
    if (!f) return 2;
    string line;
    while (getline(is, line)) {
        istringstream ls(line);
        string a, b;
        if (!(ls >> a >> b)) continue;
        set_pairwise_globals();
        MSA::SetIdCount(2);
        {
            // column scores and anchor columns on MSAs prepared as
            // AnchoredProfileProfile prepares them
            MSA m1, m2;
            msa_from_string(m1, a, 0);
            msa_from_string(m2, b, 1);
            SetAlpha(ALPHA_DNA);
            m1.FixAlpha();
            m2.FixAlpha();
            SetPPScore(PPSCORE_SPN);
            PrepareMSAforScoring(m1);
            PrepareMSAforScoring(m2);
            vector<SCORE> sc(a.size());
            LetterObjScoreXP(m1, m2, sc.data());
            fprintf(f, "S %zu", sc.size());
            for (SCORE x : sc) {
                uint32_t u;
                memcpy(&u, &x, 4);
                fprintf(f, " %08x", u);
            }
            fprintf(f, "\n");
            MSA p1, p2;
            msa_from_string(p1, a, 0);
            msa_from_string(p2, b, 1);
            p1.FixAlpha();
            p2.FixAlpha();
            PrepareMSAforScoring(p1);
            PrepareMSAforScoring(p2);
            vector<unsigned> cols(a.size() + 1);
            unsigned nc = 0;
            FindAnchorColsPP(p1, p2, cols.data(), &nc);
            fprintf(f, "A %u", nc);
            for (unsigned i = 0; i < nc; ++i) fprintf(f, " %u", cols[i]);
            fprintf(f, "\n");
        }
        set_pairwise_globals();
        MSA::SetIdCount(2);
        MSA msa1, msa2, msaOut;
        msa_from_string(msa1, a, 0);
        msa_from_string(msa2, b, 1);
        SetAlpha(ALPHA_DNA);
        msa1.FixAlpha();
        msa2.FixAlpha();
        SetPPScore(PPSCORE_SPN);
        AnchoredProfileProfile(msa1, msa2, msaOut);
        fprintf(f, "R %s %s\n", row_of(msaOut, 0).c_str(), row_of(msaOut, 1).c_str());
    }
    fclose(f);
    return 0;
}

int do_refinew(unsigned window, const char* in, const char* out)
{
    using namespace muscle;
    ifstream is(in);
    FILE* f = fopen(out, "w");
    if (!f) return 2;
    string line;
    while (getline(is, line)) {
        istringstream ls(line);
        string a, b;
        if (!(ls >> a >> b)) continue;
        set_pairwise_globals();
        g_uRefineWindow.get() = window;
        g_uWindowTo.get() = 0;
        MSA::SetIdCount(2);
        MSA msa;
        msa.SetSize(2, a.size());
        const string* rows[2] = {&a, &b};
        for (unsigned r = 0; r < 2; ++r) {
            char name[32];
            snprintf(name, sizeof(name), "seq%u", r);
            msa.SetSeqName(r, name);
            msa.SetSeqId(r, r);
            for (size_t i = 0; i < a.size(); ++i) msa.SetChar(r, i, (*rows[r])[i]);
        }
        SetAlpha(ALPHA_DNA);
        msa.FixAlpha();
        SetPPScore(PPSCORE_SPN);
        Tree GuideTree;
        TreeFromMSA(msa, GuideTree, g_Cluster2.get(), g_Distance2.get(), g_Root2.get());
        SetMuscleTree(GuideTree);
        MSA msaOut;
        RefineW(msa, msaOut);
        // RefineFast reads the rows by GetSeqIndex (LM/MuscleInterface.cpp:879-886)
        string r0, r1;
        for (unsigned r = 0; r < msaOut.GetSeqCount(); ++r) {
            const unsigned idx = msaOut.GetSeqIndex(r);
            string row(msaOut.GetSeqBuffer(idx), msaOut.GetColCount());
            (r == 0 ? r0 : r1) = row;
        }
        fprintf(f, "%s %s\n", r0.c_str(), r1.c_str());
    }
    fclose(f);
    return 0;
}

int do_nw(const char* in, const char* out)
{
    using namespace muscle;
    ifstream is(in);
    FILE* f = fopen(out, "w");
    if (!f) return 2;
    string line;
    while (getline(is, line)) {
        istringstream ls(line);
        string a, b;
        if (!(ls >> a >> b)) continue;
        g_SeqType.get() = SEQTYPE_DNA;
        g_uMaxIters.get() = 1;
        g_bStable.get() = true;
        g_bQuiet.get() = true;
        g_SeqWeight1.get() = SEQWEIGHT_ClustalW;
        SetMaxIters(g_uMaxIters.get());
        SetSeqWeightMethod(g_SeqWeight1.get());
        MSA::SetIdCount(2);
        MSA msa1, msa2, msaOut;
        msa_from_string(msa1, a, 0);
        msa_from_string(msa2, b, 1);
        SetAlpha(ALPHA_DNA);
        msa1.FixAlpha();
        msa2.FixAlpha();
        SetPPScore(PPSCORE_SPN);
        ProfileProfile(msa1, msa2, msaOut);
        // ProfileProfile leaves both ids at 0 (ProfileFromMSALocal renumbers
        // each input); rows come out as msa1's sequence then msa2's
        // (AlignTwoMSAsGivenPath), so read them by row index.
        for (unsigned r = 0; r < msaOut.GetSeqCount(); ++r) {
            const char* buf = msaOut.GetSeqBuffer(r);
            string row(buf, msaOut.GetColCount());
            fprintf(f, "%s%c", row.c_str(), r + 1 < msaOut.GetSeqCount() ? ' ' : '\n');
        }
    }
    fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc >= 5 && !strcmp(argv[1], "keys"))
        return do_keys(strtoull(argv[2], 0, 0), argv[3], argv[4]);
    if (argc >= 6 && !strcmp(argv[1], "findmer"))
        return do_findmer(strtoull(argv[2], 0, 0), argv[3], argv[4], argv[5]);
    if (argc >= 4 && !strcmp(argv[1], "app"))
        return do_app(argv[2], argv[3]);
    if (argc >= 5 && !strcmp(argv[1], "refinew"))
        return do_refinew((unsigned)strtoul(argv[2], 0, 0), argv[3], argv[4]);
    if (argc >= 4 && !strcmp(argv[1], "nw"))
        return do_nw(argv[2], argv[3]);
    fprintf(stderr, "usage: ref_harness keys <seed> <in.seq> <out.bin> | findmer <seed> <in.seq> <queries.bin> "
                    "<out.bin> | app <in.txt> <out.txt> | refinew <window> <in.txt> <out.txt> | nw <in.txt> <out.txt>\n");
    return 1;
}
