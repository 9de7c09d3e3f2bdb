"""Independent O(nm) dynamic-programming scorer (Gotoh), used to pin scores.

Maximises the alignment score exactly as the reference rescores a CIGAR
(alignment/cigar.c:277-342): +(-match) per match, -mismatch per mismatch,
and -(o + L*e) per gap run of length L; with two gap pieces a run costs
min(o1 + L*e1, o2 + L*e2).  Written from the textbook recurrence, not from
the reference code; small inputs only (pure Python).
"""
NEG = -(1 << 40)


def gotoh(a: bytes, b: bytes, match, mismatch, o1, e1, o2=None, e2=None):
    n, m = len(a), len(b)
    pieces = [(o1, e1)] + ([(o2, e2)] if o2 is not None else [])
    bonus = -match
    H_prev = [0] * (m + 1)
    E_prev = [[NEG] * (m + 1) for _ in pieces]   # gap consuming text (I)
    F_prev = [[NEG] * (m + 1) for _ in pieces]   # gap consuming pattern (D)
    for j in range(1, m + 1):
        best = NEG
        for p, (o, e) in enumerate(pieces):
            E_prev[p][j] = -(o + j * e)
            best = max(best, E_prev[p][j])
        H_prev[j] = best
    for i in range(1, n + 1):
        H = [NEG] * (m + 1)
        E = [[NEG] * (m + 1) for _ in pieces]
        F = [[NEG] * (m + 1) for _ in pieces]
        best = NEG
        for p, (o, e) in enumerate(pieces):
            F[p][0] = -(o + i * e)
            best = max(best, F[p][0])
        H[0] = best
        ai = a[i - 1]
        for j in range(1, m + 1):
            h = H_prev[j - 1] + (bonus if ai == b[j - 1] else -mismatch)
            for p, (o, e) in enumerate(pieces):
                E[p][j] = max(E[p][j - 1] - e, H[j - 1] - o - e)
                F[p][j] = max(F_prev[p][j] - e, H_prev[j] - o - e)
                h = max(h, E[p][j], F[p][j])
            H[j] = h
        H_prev, E_prev, F_prev = H, E, F
    if n == 0 and m == 0:
        return 0
    return H_prev[m]


def cigar_valid(pattern: bytes, text: bytes, ops: bytes) -> bool:
    """The CIGAR transforms pattern into text (alignment/cigar.c:347-418)."""
    i = j = 0
    for c in ops:
        if c == ord("M"):
            if i >= len(pattern) or j >= len(text) or pattern[i] != text[j]:
                return False
            i += 1
            j += 1
        elif c == ord("X"):
            if i >= len(pattern) or j >= len(text) or pattern[i] == text[j]:
                return False
            i += 1
            j += 1
        elif c == ord("I"):
            j += 1
        elif c == ord("D"):
            i += 1
        else:
            return False
    return i == len(pattern) and j == len(text)


def cigar_score(ops: bytes, match, mismatch, o1, e1, o2=None, e2=None) -> int:
    """Score of a CIGAR (gap runs priced per piece, the cheaper piece wins)."""
    score = 0
    i = 0
    n = len(ops)
    while i < n:
        j = i
        while j < n and ops[j] == ops[i]:
            j += 1
        L = j - i
        c = ops[i]
        if c == ord("M"):
            score -= match * L
        elif c == ord("X"):
            score -= mismatch * L
        else:
            cost = o1 + L * e1
            if o2 is not None:
                cost = min(cost, o2 + L * e2)
            score -= cost
        i = j
    return score


def gotoh_endsfree(a: bytes, b: bytes, match, mismatch, o1, e1, o2=None, e2=None,
                   pbf=0, pef=0, tbf=0, tef=0):
    """Best score of an alignment that may skip up to pbf / tbf leading and
    pef / tef trailing symbols of the pattern / text at no cost (semi-global
    DP: free first column / row cells, best cell of the free last row /
    column).  Written from the textbook recurrence; small inputs only."""
    n, m = len(a), len(b)
    pieces = [(o1, e1)] + ([(o2, e2)] if o2 is not None else [])
    bonus = -match
    H_prev = [0] * (m + 1)
    E_prev = [[NEG] * (m + 1) for _ in pieces]
    F_prev = [[NEG] * (m + 1) for _ in pieces]
    for j in range(1, m + 1):
        best = NEG
        for p, (o, e) in enumerate(pieces):
            E_prev[p][j] = max(E_prev[p][j - 1] - e, H_prev[j - 1] - o - e)
            best = max(best, E_prev[p][j])
        H_prev[j] = 0 if j <= tbf else best
    rows = [H_prev[:]]
    for i in range(1, n + 1):
        H = [NEG] * (m + 1)
        E = [[NEG] * (m + 1) for _ in pieces]
        F = [[NEG] * (m + 1) for _ in pieces]
        best = NEG
        for p, (o, e) in enumerate(pieces):
            F[p][0] = max(F_prev[p][0] - e, H_prev[0] - o - e)
            best = max(best, F[p][0])
        H[0] = 0 if i <= pbf else best
        ai = a[i - 1]
        for j in range(1, m + 1):
            h = H_prev[j - 1] + (bonus if ai == b[j - 1] else -mismatch)
            for p, (o, e) in enumerate(pieces):
                E[p][j] = max(E[p][j - 1] - e, H[j - 1] - o - e)
                F[p][j] = max(F_prev[p][j] - e, H_prev[j] - o - e)
                h = max(h, E[p][j], F[p][j])
            H[j] = h
        H_prev, E_prev, F_prev = H, E, F
        rows.append(H[:])
    best = max(rows[n][j] for j in range(max(0, m - tef), m + 1))
    best = max(best, max(rows[i][m] for i in range(max(0, n - pef), n + 1)))
    return best


def strip_free_ends(ops: bytes, pbf, pef, tbf, tef) -> bytes:
    """The CIGAR without the gaps its free ends absorb: up to tbf / pbf
    symbols of a leading I / D run and up to tef / pef of a trailing one."""
    def lead(s, c, k):
        i = 0
        while i < len(s) and i < k and s[i] == c:
            i += 1
        return s[i:]
    ops = lead(ops, ord("I"), tbf) if ops[:1] == b"I" else lead(ops, ord("D"), pbf)
    r = ops[::-1]
    r = lead(r, ord("I"), tef) if r[:1] == b"I" else lead(r, ord("D"), pef)
    return r[::-1]


def maxtrim(ops: bytes, metric, match, mismatch, o1, e1=0, o2=None, e2=None):
    """Longest-prefix-of-best-score trim of an extension CIGAR, from its
    definition: a prefix's score counts 1 per match when match is 0 (else
    -match), -mismatch per mismatch and gap runs as in cigar_score; for
    dual-cost gaps the score is only taken at run ends.  Returns
    (kept_ops, score, trimmed) with score None when nothing scores above 0."""
    m = -match if match != 0 else 1
    best, best_len, score = 0, 0, 0
    runs = []
    for c in ops:
        if runs and runs[-1][0] == c:
            runs[-1][1] += 1
        else:
            runs.append([c, 1])
    pos = 0
    for c, L in runs:
        for t in range(1, L + 1):
            if metric == "affine2p" and t < L:
                continue
            if c == ord("M"):
                s = score + m * t
            elif c == ord("X"):
                s = score - mismatch * t
            elif metric == "linear":
                s = score - o1 * t
            elif metric == "affine":
                s = score - (o1 + e1 * t)
            else:
                s = score - min(o1 + e1 * t, o2 + e2 * t)
            if s > best:
                best, best_len = s, pos + t
        if c == ord("M"):
            score += m * L
        elif c == ord("X"):
            score -= mismatch * L
        elif metric == "linear":
            score -= o1 * L
        elif metric == "affine":
            score -= o1 + e1 * L
        else:
            score -= min(o1 + e1 * L, o2 + e2 * L)
        pos += L
    trimmed = best_len != len(ops)
    if metric == "affine2p" and not ops:
        trimmed = False
    if best == 0:
        return b"", None, trimmed
    return ops[:best_len], best, trimmed
