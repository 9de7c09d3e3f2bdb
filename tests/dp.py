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
