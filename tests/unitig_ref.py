"""Reference maximal unitigs for small inputs, computed from global brute-force
side counts (every (k+1)-mer instance of every fragment, no partitioning), with
the reference's walk (include/Subgraph.hpp:414-543) and output orientation
(include/Unitig_Collator.hpp:306-400).  Vertices in no edge at the cutoff are
not part of the graph and yield no unitig.  Test infrastructure."""
import oracle_lib as O

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def side_counts(seq, ro, k):
    beg, ln = O.fragments(seq, ro, k)
    cnt = {}
    verts = set()
    for b, n in zip(beg, ln):
        frag = seq[int(b):int(b) + int(n)].decode().upper()
        for i in range(len(frag) - k + 1):
            x = frag[i:i + k]
            verts.add(min(x, rc(x)))
        for i in range(len(frag) - k):
            u, v = frag[i:i + k], frag[i + 1:i + k + 1]
            uc, vc = min(u, rc(u)), min(v, rc(v))
            succ, pred = frag[i + k], frag[i]
            key = (uc, 1, succ) if u == uc else (uc, 0, COMP[succ])
            cnt[key] = cnt.get(key, 0) + 1
            if uc != vc:
                key = (vc, 0, pred) if v == vc else (vc, 1, COMP[pred])
                cnt[key] = cnt.get(key, 0) + 1
    return cnt, verts


def edges(cnt, v, side, f_th):
    return [b for b in "ACGT" if min(15, cnt.get((v, side, b), 0)) >= f_th]


def canonical_linear(u, k):
    r = rc(u)
    for t in range(k):
        if u[t] < r[t]:
            return u
        if u[t] > r[t]:
            return r
    return u


def canonical_cycle(lab, k):
    n = len(lab) - (k - 1)
    cyc = lab[:n]
    best = None
    for c in (cyc, rc(cyc)):
        for st in range(n):
            r = c[st:] + c[:st]
            r = (r * (1 + (k - 1 + n - 1) // n))[: n + k - 1]  # periodic: cycles shorter than k - 1 too
            if best is None or r[:k] < best[:k] or (r[:k] == best[:k] and r < best):
                best = r
    return best


def unitigs(seq, ro, k, f_th):
    cnt, verts = side_counts(seq, ro, k)
    order = sorted(verts)
    visited = set()

    def walk(v_hat, s0):
        x = v_hat if s0 == 1 else rc(v_hat)
        s = s0
        cur = v_hat
        ext = []
        while True:
            e = edges(cnt, cur, s, f_th)
            if len(e) != 1:
                return ext, False
            b = e[0] if s == 1 else COMP[e[0]]
            x = x[1:] + b
            u = min(x, rc(x))
            ent = 0 if x == u else 1
            if len(edges(cnt, u, ent, f_th)) > 1:
                return ext, False
            if u in visited:
                return ext, (u == v_hat and ent == 1 - s0)
            visited.add(u)
            ext.append(b)
            cur, s = u, 1 - ent

    out = []
    for v in order:
        if v in visited:
            continue
        if not edges(cnt, v, 0, f_th) and not edges(cnt, v, 1, f_th):
            continue  # isolated: in no edge at the cutoff, never output (src/Subgraph.cpp:206-214)
        visited.add(v)
        back, cyc = walk(v, 1)
        if cyc:
            out.append(canonical_cycle(v + "".join(back), k))
            continue
        front, _ = walk(v, 0)
        out.append(canonical_linear(rc("".join(front)) + v + "".join(back), k))
    return sorted(out)
