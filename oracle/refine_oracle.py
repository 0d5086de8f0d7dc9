"""CPU oracle of the refinement of a pairwise gapped alignment: the job
producers around the DP and the splicing of its paths.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
mh_refine_* entry points; nothing in the product calls it.

Restates, for two genomes (single-sequence profiles, sequence weight 1):
  removeLargeGapsPP      LM/ProgressiveAligner.cpp:1051-1117 (split where more
                         than 5000 columns are not aligned in both rows, keeping
                         300 columns of hangover; those pieces are not refined)
  IsDenseEnough          LM/ProgressiveAligner.cpp:1015-1038 and the halving
                         loop of refineAlignment 1142-1166
  ProfileAlignFast       LM/MuscleInterface.cpp:1053-1162 with anchored=true:
  AnchoredProfileProfile MU/anchoredpp.cpp:443-552
    LetterObjScoreXP     MU/anchoredpp.cpp:16-316 (per-column letter score
                         NUC_SP, and gap penalties spread over the gap's columns)
    WindowSmooth         MU/anchors.cpp:9-48 (window 21)
    FindBestColsComboPP  MU/anchoredpp.cpp:327-350 (score and smoothed score >= 90)
    MergeBestCols        MU/anchors.cpp:137-186 (spacing 96)
    ColsToRanges         MU/refinevert.cpp:32-50
  RefineFast / RefineW   LM/MuscleInterface.cpp:823-894, MU/refinew.cpp:75-200:
                         windows of 500 or 200 columns re-aligned from their
                         residues (MUSCLE() on two sequences is one NWSmall of
                         the first against the second, MU/muscle.cpp:85-102,
                         MU/progressivealign.cpp:14-80); window size 0 is
                         RefineVert, which returns at once for fewer than three
                         sequences (MU/refinevert.cpp:60-61)
  refineAlignment        LM/ProgressiveAligner.cpp:1119-1240 (dispatch and
                         reassembly), run as progressiveMauve runs it for two
                         genomes: the profile pass, then the final refinement
                         pass (LM/ProgressiveAligner.cpp:2561-2565, 3751-3755)
Scores are float32, operation by operation, as MUSCLE's SCORE is float.

Reproduced as the reference does it: a piece in which one genome has no
residues is re-aligned in the refinement pass from its one row, and the
reassembly then writes that row into genome 1's row and does not advance
genome 2's (LM/ProgressiveAligner.cpp:1222-1237 with
LM/MuscleInterface.cpp:826-833, 887): genome 2's later pieces move left.

Parity pins (tests/test_refine_oracle.py): hand-worked column scores (gap
penalties spread per column), the smoothing against direct window sums on
integer scores, MergeBestCols cases, and the DP pieces against the oracle
NWSmall (the DP itself is pinned to libMUSCLE through oracle/mh_oracle.c).
"""
import numpy as np

from oracle import pyoracle as po

F = np.float32
GAP_OPEN = F(-400)
GAP_EXTEND = F(0)
TERM_GAP = GAP_OPEN / F(2)                  # TERMGAPS_Half, MU/objscore2.cpp:22-41
_HOXD = ((91, -114, -31, -123), (-114, 100, -125, -31), (-31, -125, 100, -114), (-123, -31, -114, 91))
NUC_SP = [[F(x) + F(60) for x in row] for row in _HOXD]     # MU/nucmx.cpp:8-25
SMOOTH_WINDOW = 21
ANCHOR_SPACING = 96
MIN_BEST_COL_SCORE = 90.0
MIN_SMOOTH_SCORE = 90.0
SMOOTH_CEIL = 999.0
MAX_GAP_LENGTH = 5000
LCB_HANGOVER = 300
MIN_WINDOW = 200.0
MAX_WINDOW = 20000.0
MIN_DENSITY = 0.5
MAX_DENSITY = 0.9
INSANE = 0xFFFFFFFF


def _mu_gap(c):
    return c == "-" or c == "."


def _letter(c):
    """Letter index of a residue for the column score: A,C,G,T (U as T, as
    libMUSCLE scores it, tests/golden/refine_golden.txt), else 99 (no letter
    score: wildcards)."""
    i = "ACGTU".find(c.upper())
    return min(i, 3) if i >= 0 else 99


def column_scores(r1, r2):
    """LetterObjScoreXP's per-column score for one pair of rows."""
    n = len(r1)
    mm = [F(0)] * n
    gg = [F(0)] * n
    start, end = 0, n - 1
    for c in range(n):
        if not (_mu_gap(r1[c]) and _mu_gap(r2[c])):
            start = c
            break
    for c in range(n - 1, -1, -1):
        if not (_mu_gap(r1[c]) and _mu_gap(r2[c])):
            end = c
            break
    for c in range(start, end + 1):
        a = _letter(r1[c])
        if a >= 4:
            continue
        b = _letter(r2[c])
        if b >= 4:
            continue
        mm[c] = NUC_SP[a][b]
    g1 = g2 = False
    left = 0
    cur = F(0)
    for c in range(start, end + 1):
        x, y = _mu_gap(r1[c]), _mu_gap(r2[c])
        if x and y:
            continue
        if x or y:
            if (x and not g1) or (not x and not g2):
                left = c
                cur = F(cur + (TERM_GAP if c == start else GAP_OPEN))
                if x:
                    g1 = True
                else:
                    g2 = True
            else:
                cur = F(cur + GAP_EXTEND)
            continue
        if g1 or g2:
            per = F(cur / F(c - left))
            for k in range(left, c):
                gg[k] = per
            left = INSANE
            cur = F(0)
        g1 = g2 = False
    if g1 or g2:
        cur = F(cur - GAP_OPEN)
        cur = F(cur + TERM_GAP)
        per = F(cur / F(n - left))
        for k in range(left, n):
            gg[k] = per
    return [F(F(0) + F(F(1) * F(mm[c] + gg[c]))) for c in range(n)]


def window_smooth(score, w=SMOOTH_WINDOW, ceil=SMOOTH_CEIL):
    n = len(score)
    out = [F(0)] * n
    if n <= w:
        return out
    h = w // 2

    def cl(x):
        return F(ceil) if float(x) > ceil else F(x)
    tot = F(0)
    for i in range(w):
        tot = F(tot + cl(score[i]))
    i = h
    while True:
        out[i] = F(tot / F(w))
        if i == n - h - 1:
            break
        tot = F(tot - cl(score[i - h]))
        tot = F(tot + cl(score[i + h + 1]))
        i += 1
    return out


def best_cols(score, smooth, min_score=MIN_BEST_COL_SCORE, min_smooth=MIN_SMOOTH_SCORE):
    return [i for i in range(len(score)) if not (float(score[i]) < min_score) and not (float(smooth[i]) < min_smooth)]


def merge_best_cols(score, best, spacing=ANCHOR_SPACING):
    out = []
    n = 0
    while n < len(best):
        b = best[n]
        within = 0
        for i in range(n + 1, len(best)):
            if best[i] - b >= spacing:
                break
            within += 1
        anchor = b
        if within == 1:
            b2 = best[n + 1]
            anchor = b if score[b] > score[b2] else b2
        elif within > 1:
            closest_d, closest = spacing, b
            for i in range(n + 1, n + within):
                d = abs(best[i] - b)
                if d < closest_d:
                    closest, closest_d = best[i], d
            anchor = closest
        out.append(anchor)
        n += within + 1
    return out


def cols_to_ranges(anchors, n):
    return [(anchors[k - 1] if k > 0 else 0, anchors[k] if k < len(anchors) else n) for k in range(len(anchors) + 1)]


def anchor_cols(r1, r2):
    s = column_scores(r1, r2)
    return merge_best_cols(s, best_cols(s, window_smooth(s)))


def _residues(row):
    return "".join(c for c in row if not _mu_gap(c))


def nwsmall_rows(a, b):
    """The NWSmall alignment of residue strings a (as A) and b (as B)."""
    if not a or not b:
        raise ValueError("NWSmall: length=0")
    path, _ = po.nwsmall(a.upper(), b.upper())
    return po.rows_from_path(a, b, path)


def anchored_profile_profile(r1, r2, dp=nwsmall_rows):
    o1, o2 = [], []
    for lo, hi in cols_to_ranges(anchor_cols(r1, r2), len(r1)):
        if hi == lo:
            continue
        x, y = dp(_residues(r1[lo:hi]), _residues(r2[lo:hi]))
        o1.append(x)
        o2.append(y)
    return "".join(o1), "".join(o2)


def refine_w(rows, window, dp=nwsmall_rows):
    """RefineW over the rows of the genomes present (one or two)."""
    n = len(rows[0])
    out = [[] for _ in rows]
    for wi in range((n + window - 1) // window):
        f = wi * window
        t = min(f + window - 1, n - 1)
        seqs = [_residues(r[f:t + 1]) for r in rows]
        full = [k for k, s in enumerate(seqs) if s]
        if len(full) > 1:
            x, y = dp(seqs[0], seqs[1])
            out[0].append(x)
            out[1].append(y)
        elif len(full) == 1:
            k = full[0]
            for j in range(len(rows)):
                out[j].append(seqs[k] if j == k else "-" * len(seqs[k]))
        else:
            raise ValueError("window without residues")
    return ["".join(o) for o in out]


def remove_large_gaps(r1, r2):
    """[(first col, end col, unaligned piece)] of removeLargeGapsPP."""
    n = len(r1)
    pieces = []
    gap_cols = 0
    last = None
    base = 0
    for col in range(n):
        if r1[col] != "-" and r2[col] != "-":
            if gap_cols > MAX_GAP_LENGTH:
                if last is not None:
                    sp = last + LCB_HANGOVER - base
                    pieces.append((base, base + sp, False))
                    base += sp
                sp = col - LCB_HANGOVER - base
                pieces.append((base, base + sp, True))
                base += sp
            last = col
            gap_cols = 0
        else:
            gap_cols += 1
    if gap_cols > MAX_GAP_LENGTH:
        if last is not None:
            sp = last + LCB_HANGOVER - base
            pieces.append((base, base + sp, False))
            base += sp
        pieces.append((base, n, True))
    else:
        pieces.append((base, n, False))
    return pieces


def _lengths(r1, r2, lo, hi):
    return (sum(1 for c in r1[lo:hi] if c != "-"), sum(1 for c in r2[lo:hi] if c != "-"))


def dense_enough(r1, r2, lo, hi):
    l1, l2 = _lengths(r1, r2, lo, hi)
    mult = (l1 > 0) + (l2 > 0)
    aln = hi - lo
    density = float(l1 + l2) / (aln * float(mult))
    threshold = ((MAX_DENSITY - MIN_DENSITY) / (MIN_WINDOW - MAX_WINDOW)) * (float(aln) - MAX_WINDOW) + MIN_DENSITY
    if density > MAX_DENSITY:
        return 2
    if density > threshold:
        return 1
    return 0


def pieces_of(r1, r2):
    """removeLargeGapsPP, then the halving of pieces too long for their density."""
    pieces = remove_large_gaps(r1, r2)
    i = 0
    while i < len(pieces):
        lo, hi, g = pieces[i]
        d = dense_enough(r1, r2, lo, hi)
        n = hi - lo
        if (d == 0 and n > MAX_WINDOW / 3) or (d == 1 and n > MAX_WINDOW) or (d == 2 and n > MAX_WINDOW * 3):
            sp = n // 2
            pieces[i] = (lo, hi - sp, g)
            pieces.insert(i + 1, (hi - sp, hi, g))
            continue
        i += 1
    return pieces


def refine_alignment(r1, r2, profile_aln, dp=nwsmall_rows):
    outs = []
    for lo, hi, gap in pieces_of(r1, r2):
        a, b = r1[lo:hi], r2[lo:hi]
        l1, l2 = _lengths(r1, r2, lo, hi)
        if gap:
            outs.append([a, b])
        elif profile_aln:
            outs.append(list(anchored_profile_profile(a, b, dp)) if l1 > 0 and l2 > 0 else [a, b])
        else:
            d = dense_enough(r1, r2, lo, hi)
            present = [r for r, ln in ((a, l1), (b, l2)) if ln > 0]
            outs.append(present if d == 0 else refine_w(present, 500 if d == 1 else 200, dp))
    new_len = sum(len(o[0]) for o in outs)
    rows = [["-"] * new_len, ["-"] * new_len]
    pos = [0, 0]
    for o in outs:
        for k, r in enumerate(o):
            rows[k][pos[k]:pos[k] + len(r)] = list(r)
            pos[k] += len(r)
    return "".join(rows[0][:new_len]), "".join(rows[1][:new_len])


def gapped_alignment(r1, r2, dp=nwsmall_rows):
    """The profile pass, then the final refinement pass."""
    a, b = refine_alignment(r1, r2, True, dp)
    return refine_alignment(a, b, False, dp)
