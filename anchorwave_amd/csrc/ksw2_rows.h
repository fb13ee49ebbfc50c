// ksw2_rows.h -- AnchorWave's host-side handling of a ksw_extd2 result:
// its code table, the gapped rows a CIGAR spells, their two-piece affine
// rescoring and the reconstruction check
// (src/myImportandFunction/alignSlidingWindow.cpp:644-721, :852-870).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <string>

namespace kswhip {

// alignSlidingWindow.cpp:644-660 (and the adapters' table, :820-838).
inline uint8_t code_of(char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

// Appends the rows a CIGAR spells from positions (qp, tp) of the two
// sequences (:852-870; std::string::append clamps at the sequence end) and
// advances them.
inline void append_rows(std::string &aq, std::string &ad, const char *dq, int64_t ql, int64_t &qp, const char *dd,
                        int64_t tl, int64_t &tp, const uint32_t *cigar, int n)
{
    auto take = [](std::string &row, const char *seq, int64_t len, int64_t pos, int64_t count) {
        if (pos < len) row.append(seq + pos, (size_t)std::min(count, len - pos));
    };
    for (int i = 0; i < n; ++i) {
        const int64_t len = cigar[i] >> 4;
        const int op = (int)(cigar[i] & 0xf);
        if (op == 0 || op == 7 || op == 8) {
            take(aq, dq, ql, qp, len), take(ad, dd, tl, tp, len);
            qp += len, tp += len;
        } else if (op == 1) {
            take(aq, dq, ql, qp, len), ad.append((size_t)len, '-');
            qp += len;
        } else if (op == 2) {
            aq.append((size_t)len, '-'), take(ad, dd, tl, tp, len);
            tp += len;
        }
    }
}

// gappedAlignmentReconstructs (:662-675).
inline bool reconstructs(const std::string &row, const char *seq, int64_t len)
{
    int64_t pos = 0;
    for (const char c : row) {
        if (c == '-') continue;
        if (pos >= len || seq[pos] != c) return false;
        ++pos;
    }
    return pos == len;
}

// scoreGappedTwoPieceAffineAlignment (:677-721); *ok = false where the
// reference throws (rows of different lengths, a column of two gaps).
inline int64_t rescore_rows(const std::string &aq, const std::string &ad, int32_t mis, int32_t o1, int32_t e1,
                            int32_t o2, int32_t e2, bool *ok)
{
    *ok = aq.size() == ad.size();
    if (!*ok) return 0;
    int64_t score = 0;
    size_t col = 0;
    while (col < aq.size()) {
        const bool qgap = aq[col] == '-', dgap = ad[col] == '-';
        if (qgap && dgap) {
            *ok = false;
            return score;
        }
        if (qgap || dgap) {
            size_t end = col + 1;
            while (end < aq.size() && (aq[end] == '-') == qgap && (ad[end] == '-') != qgap) ++end;
            const int64_t l = (int64_t)(end - col);
            score += std::max<int64_t>((int64_t)o1 + (int64_t)e1 * l, (int64_t)o2 + (int64_t)e2 * l);
            col = end;
            continue;
        }
        if (aq[col] != ad[col]) score += mis;
        ++col;
    }
    return score;
}

}  // namespace kswhip
