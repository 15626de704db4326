"""CPU emulation of the load-balanced product stream of the ParILUT row
kernels (pt_stream in xsolver_amd/csrc/xs_parilut.hip): for every wave step
and lane, the part found by the ballot bounds and the uniform loop over the
parts the step spans, and the position q = base_part + p, must enumerate
exactly the pairs (L entry, U-part entry) of the row, each once, whatever
the part lengths (empty parts, parts longer than a step, more than 64 L
entries)."""
import numpy as np
import pytest

WAVE = 64
P = 2          # positions per lane per step (PT_SPOS)


def stream(lb, le, off, frp, fci, fdpos):
    """pairs (L position c, U position q) in the order the kernel visits them"""
    out = []
    for c0 in range(lb, le, WAVE):
        cnt = min(WAVE, le - c0)
        ub = np.zeros(WAVE, np.int64)
        ln = np.zeros(WAVE, np.int64)
        for lane in range(cnt):
            k = fci[c0 + lane]
            ub[lane] = fdpos[k] + off
            ln[lane] = frp[k + 1] - ub[lane]
        inc = np.cumsum(ln)
        tot = int(inc[-1])
        st = inc - ln
        base = ub - st
        for p0 in range(0, tot, P * WAVE):
            pe = min(p0 + P * WAVE, tot) - 1
            o0 = int(np.sum(inc <= p0))          # popcount of a ballot
            o1 = int(np.sum(inc <= pe))
            assert 0 <= o0 <= o1 < cnt
            for lane in range(WAVE):
                for p in (p0 + c * WAVE + lane for c in range(P)):
                    q = owner = None
                    for t in range(o0, o1 + 1):
                        if p >= st[t]:
                            q, owner = base[t] + p, t
                    if p < tot:
                        assert owner is not None and ln[owner] > 0
                        assert st[owner] <= p < inc[owner]
                        out.append((c0 + owner, int(q)))
    return out


def random_factor(rng, n, maxlen):
    """combined-factor rows: sorted columns with the pivot, random fill"""
    frp = [0]
    fci = []
    fdpos = []
    for i in range(n):
        cols = set(rng.choice(n, size=rng.integers(0, maxlen), replace=False).tolist())
        cols.add(i)
        cols = sorted(cols)
        fdpos.append(frp[-1] + cols.index(i))
        fci += cols
        frp.append(len(fci))
    return np.array(frp), np.array(fci), np.array(fdpos)


@pytest.mark.parametrize("maxlen,off", [(3, 1), (40, 1), (150, 0), (300, 1)])
def test_stream_enumerates_every_product_once(maxlen, off):
    rng = np.random.default_rng(maxlen)
    n = 400
    frp, fci, fdpos = random_factor(rng, n, maxlen)
    for i in list(range(0, n, 37)) + [n - 1]:
        lb, le = frp[i], fdpos[i]
        got = stream(lb, le, off, frp, fci, fdpos)
        want = [(c, q) for c in range(lb, le) for q in range(fdpos[fci[c]] + off, frp[fci[c] + 1])]
        assert sorted(got) == sorted(want)
        assert len(got) == len(set(got))


def test_stream_more_than_64_l_entries_and_empty_parts():
    # a row with 130 L entries whose rows k alternate between an empty
    # strict upper part and one longer than a wave step
    n = 400
    frp = [0]
    fci = []
    fdpos = []
    for k in range(n):
        if k < 300:
            cols = [k] if k % 2 else [k] + list(range(k + 1, min(n, k + 1 + 140)))
        else:
            cols = list(range(0, 130)) + [k] if k == 399 else [k]
        fdpos.append(len(fci) + cols.index(k))
        fci += cols
        frp.append(len(fci))
    frp, fci, fdpos = np.array(frp), np.array(fci), np.array(fdpos)
    i = 399
    got = stream(frp[i], fdpos[i], 1, frp, fci, fdpos)
    want = [(c, q) for c in range(frp[i], fdpos[i]) for q in range(fdpos[fci[c]] + 1, frp[fci[c] + 1])]
    assert sorted(got) == sorted(want) and len(got) == len(want)
