"""Checkers and input recipes that run where the arrays are (torch tensors on
the GPU for the full-size tests; the same code on the CPU in the checkers'
own tests).

Arrays of 1 GiB and more are checked without leaving the device.  A sort's
output is fixed, up to the order inside the +-0 run and the NaN payloads, by
two properties that the reference's checkers also rest on
(tests/test-qsort-common.h:29-50): it is ordered under the reference's
comparator (utils/custom-compare.h:11-34) and it is a permutation of the
input.  Both are tested at once: the output's order keys must equal the
input's order keys sorted by torch.sort.  The order key maps every NaN to one
class above +inf (first when descending) and -0.0 to +0.0, so the remaining
bit-level facts are checked separately: the count of -0.0 bit patterns, and
qsort(hasnan)'s canonical NaN (src/xss-custom-float.h:34-38,81-85).

Keys travel as signed integer tensors of the key's width holding its bits
(int16 / int32 / int64), the form torch can sort.
"""
import numpy as np
import torch

WIDTH = {"u16": 16, "i16": 16, "f16": 16, "u32": 32, "i32": 32, "f32": 32, "u64": 64, "i64": 64, "f64": 64}
SIGNED = {16: torch.int16, 32: torch.int32, 64: torch.int64}
FLOATS = ("f16", "f32", "f64")
EXP = {16: 0x7C00, 32: 0x7F800000, 64: 0x7FF0000000000000}
QNAN = {16: 0x7C01, 32: 0x7FC00000, 64: 0x7FF8000000000000}


def _mn(nb):   # bit pattern of the sign bit alone (-0.0 for floats) as a signed value
    return -(1 << (nb - 1))


def _mag(nb):  # all bits but the sign
    return (1 << (nb - 1)) - 1


def order_key(x, tag, desc=False):
    """Signed integer tensor whose order is the reference's comparator order
    for keys of type `tag` held as bits in `x`: every NaN one class above
    +inf, -0.0 == +0.0; complemented for descending (NaNs first)."""
    nb = WIDTH[tag]
    if tag in FLOATS:
        b = torch.where(x == _mn(nb), torch.zeros_like(x), x)
        nan = (b & _mag(nb)) > EXP[nb]
        s = torch.where(b < 0, b ^ _mag(nb), b)
        s = torch.where(nan, torch.full_like(s, _mag(nb)), s)
    elif tag.startswith("u"):
        s = x ^ _mn(nb)
    else:
        s = x.clone()
    return ~s if desc else s


def from_order_key(s, tag):
    """Inverse of order_key (ascending) for keys that are not NaN."""
    nb = WIDTH[tag]
    if tag in FLOATS:
        return torch.where(s < 0, s ^ _mag(nb), s)
    if tag.startswith("u"):
        return s ^ _mn(nb)
    return s.clone()


def is_nan(x, tag):
    nb = WIDTH[tag]
    if tag not in FLOATS:
        return torch.zeros(x.shape, dtype=torch.bool, device=x.device)
    return (x & _mag(nb)) > EXP[nb]


def gen(tag, recipe, n, seed, device="cuda"):
    """The five full-size recipes, drawn on `device`: uniform over the full
    bit range (finite for floats), 1024 distinct values, sorted, reverse
    sorted, and 1% canonical NaN at random positions (floats); and "unit",
    floats uniform in (-1, 1), whose skewed exponents leave buckets too
    large for LDS after the two MSD digits."""
    nb = WIDTH[tag]
    g = torch.Generator(device=device)
    g.manual_seed(seed)

    def bits(m):
        if nb == 64:
            hi = torch.randint(-2 ** 31, 2 ** 31, (m,), dtype=torch.int64, device=device, generator=g)
            lo = torch.randint(0, 2 ** 32, (m,), dtype=torch.int64, device=device, generator=g)
            x = (hi << 32) | lo
        else:
            x = torch.randint(-2 ** (nb - 1), 2 ** (nb - 1), (m,), dtype=SIGNED[nb], device=device, generator=g)
        if tag in FLOATS:  # exponent all ones (inf / NaN): clear its top bit
            x = torch.where((x & EXP[nb]) == EXP[nb], x ^ (1 << (nb - 2)), x)
        return x

    if recipe == "random":
        return bits(n)
    if recipe == "dups":
        pool = bits(1024)
        return pool[torch.randint(0, 1024, (n,), device=device, generator=g)]
    if recipe in ("sorted", "reverse"):
        s = torch.sort(order_key(bits(n), tag)).values
        x = from_order_key(s, tag)
        return x.flip(0) if recipe == "reverse" else x
    if recipe == "unit":
        ft = {16: torch.float16, 32: torch.float32, 64: torch.float64}[nb]
        u = torch.rand(n, device=device, generator=g, dtype=torch.float64 if nb == 64 else torch.float32)
        return (u * 2 - 1).to(ft).view(SIGNED[nb])
    if recipe == "nan":
        x = bits(n)
        if tag in FLOATS:
            x[torch.rand(n, device=device, generator=g) < 0.01] = QNAN[nb]
        return x
    raise ValueError(recipe)


SORT_MAX = (1 << 31) - 1  # torch sorts and indexes at most INT_MAX elements at once
CHUNK = 1 << 30


def _counts16(o):
    """Histogram of 16-bit order keys (65536 bins), chunk by chunk."""
    c = torch.zeros(1 << 16, dtype=torch.int64, device=o.device)
    for i in range(0, o.numel(), CHUNK):
        c += torch.bincount(o[i:i + CHUNK].to(torch.int32) + (1 << 15), minlength=1 << 16)
    return c


def gather(x, idx):
    """x[idx] for index tensors of any length."""
    if idx.numel() <= SORT_MAX:
        return x[idx]
    return torch.cat([x[idx[i:i + CHUNK]] for i in range(0, idx.numel(), CHUNK)])


def check_sorted_perm(got, inp, tag, desc, what):
    """got is inp ordered by the reference's comparator (value classes)."""
    og = order_key(got, tag, desc)
    if og.numel() > SORT_MAX:
        # beyond one torch.sort: 16-bit keys, checked as ordered plus equal
        # counts of every order key
        assert WIDTH[tag] == 16, "arrays over 2^30 keys are checked for 16-bit keys only"
        for i in range(0, og.numel() - 1, CHUNK):
            seg = og[i:i + CHUNK + 1]
            if not bool((seg[1:] >= seg[:-1]).all()):
                raise AssertionError(f"{what}: not ordered near {i}")
        if not torch.equal(_counts16(og), _counts16(order_key(inp, tag, desc))):
            raise AssertionError(f"{what}: not a permutation of the input")
        return
    oi = torch.sort(order_key(inp, tag, desc)).values
    if not torch.equal(og, oi):
        bad = torch.nonzero(og != oi)[0].item()
        raise AssertionError(f"{what}: order keys differ first at {bad} (n={got.numel()}): "
                             f"got {og[max(0, bad - 2):bad + 3].tolist()} expected {oi[max(0, bad - 2):bad + 3].tolist()}")


def check_sort(got, inp, tag, hasnan, desc, canon_nan, what):
    """A qsort result: ordered permutation, the same count of -0.0 bit
    patterns and, for qsort(hasnan), only canonical NaNs."""
    check_sorted_perm(got, inp, tag, desc, what)
    if tag in FLOATS:
        nb = WIDTH[tag]
        assert int((got == _mn(nb)).sum()) == int((inp == _mn(nb)).sum()), f"{what}: -0.0 count differs"
        if hasnan and canon_nan:
            assert bool((got[is_nan(got, tag)] == QNAN[nb]).all()), f"{what}: NaN not canonical"


def check_permutation(idx, n, what):
    """idx (int64) holds every index 0..n-1 once."""
    assert idx.numel() == n, f"{what}: {idx.numel()} indices for {n} keys"
    assert bool((idx >= 0).all()) and bool((idx < n).all()), f"{what}: index out of range"
    seen = torch.zeros(n, dtype=torch.int32, device=idx.device)
    seen.index_add_(0, idx, torch.ones_like(idx, dtype=torch.int32))
    assert bool((seen == 1).all()), f"{what}: indices are not a permutation"


def check_gather(keys_out, inp, idx, what):
    """idx is a permutation and inp[idx] is keys_out, bit for bit."""
    check_permutation(idx, inp.numel(), what)
    assert torch.equal(inp[idx], keys_out), f"{what}: indices do not gather the keys"


def seg_ids(offsets, device):
    """Segment id of every position of back-to-back segments (offsets: n+1)."""
    off = torch.as_tensor(np.asarray(offsets, dtype=np.int64), device=device)
    lens = off[1:] - off[:-1]
    return torch.repeat_interleave(torch.arange(lens.numel(), device=device), lens)


def check_segments_sorted_perm(got, inp, seg, tag, desc, what):
    """Each segment of got is its segment of inp ordered by the comparator:
    (segment, order key) pairs, sorted, must equal those of the output in
    place.  seg: the segment id of every position."""
    nb = WIDTH[tag]

    def comp(x):
        o = order_key(x, tag, desc).to(torch.int64)
        if nb < 64:
            return (seg.to(torch.int64) << nb) | (o - _mn(nb))  # non-negative order key in the low bits
        raise NotImplementedError("64-bit keys: composite would overflow")

    cg, ci = comp(got), torch.sort(comp(inp)).values
    if not torch.equal(cg, ci):
        bad = torch.nonzero(cg != ci)[0].item()
        raise AssertionError(f"{what}: first mismatch at {bad} (segment {int(seg[bad])})")
