"""Race emulation of the checkerboarded Metropolis schedule, on the CPU.

Within one sublattice pass every thread writes the links of its own site and
reads its staple neighbourhood (GaugeField.hpp:187-237): U_nu(x+mu),
U_mu(x+nu), U_nu(x), U_nu(x+mu-nu), U_mu(x-nu), U_nu(x-nu).  The pass is
race free when no link read on behalf of one site is written by another
site of the same pass (the reference relies on this without checking it,
Metropolis.hpp:127-163).  The sets are built from the schedule the oracle
records (the pass that updates each site, the same map the GPU kernel's own
dump equals bit for bit in the GPU tests), so the check covers the schedule
the product runs.  With an odd extent the periodic wrap puts neighbours in
the same pass, which is why the library rejects odd extents.
"""
import itertools

import numpy as np
import pytest

import oracle as orc


def schedule(nd, dims):
    """pass index of every site (lexicographic), as the oracle's sweep
    records it"""
    U = orc.identity(nd, 1, dims)
    _, _, visit = orc.metropolis_sweep(U, nd, 1, dims, 0, 1.0, 0.1, 1, 0, visit=True)
    return visit.reshape(dims)


def pass_conflicts(nd, dims, visit):
    """number of (reader site, writer site, link) triples within one pass
    where the writer is another site"""
    sites = np.array(list(itertools.product(*[range(L) for L in dims])))
    L = np.array(dims)

    def lex(x):
        return np.ravel_multi_index(tuple((x % L).T), dims)

    conflicts = 0
    for p in range(1 << nd):
        xs = sites[visit.reshape(-1) == p]
        writer = np.full(int(np.prod(dims)), -1)
        writer[lex(xs)] = lex(xs)  # a site writes its own nd links
        for mu in range(nd):
            for nu in range(nd):
                if nu == mu:
                    continue
                e_mu, e_nu = np.eye(nd, dtype=int)[mu], np.eye(nd, dtype=int)[nu]
                for off in (e_mu, e_nu, 0 * e_mu, e_mu - e_nu, -e_nu):
                    y = lex(xs + off)
                    w = writer[y]
                    conflicts += int(np.sum((w >= 0) & (w != lex(xs))))
    return conflicts


@pytest.mark.parametrize("nd,dims", [(2, (8, 6)), (3, (4, 6, 8)), (4, (4, 4, 6, 4)), (4, (8, 8, 8, 8))])
def test_sublattice_passes_are_race_free(nd, dims):
    visit = schedule(nd, dims)
    # every site is updated exactly once per sweep, by one of the 2^nd passes
    assert visit.min() == 0 and visit.max() == (1 << nd) - 1
    counts = np.bincount(visit.reshape(-1), minlength=1 << nd)
    assert np.all(counts == int(np.prod(dims)) // (1 << nd))
    assert pass_conflicts(nd, dims, visit) == 0


def test_odd_extent_would_race():
    """The same parity classes on an odd extent (the reference's schedule
    skips the last slice instead, Q6): x and x + mu with x_mu = L - 1 wrap to
    x_mu = 0 with the same parity, so one pass reads what another of its
    sites writes."""
    nd, dims = 2, (5, 4)
    x = np.array(list(itertools.product(range(5), range(4))))
    visit = ((x[:, 0] & 1) << 1 | (x[:, 1] & 1)).reshape(dims)
    assert pass_conflicts(nd, dims, visit) > 0


def test_two_colour_checkerboard_would_race():
    """An even/odd schedule (two passes by total parity) is not enough for the
    staple: x + mu - nu has the parity of x, so 2^nd parity classes are
    needed (oddeven_array, IndexHelper.hpp:84-93)."""
    nd, dims = 4, (4, 4, 4, 4)
    x = np.array(list(itertools.product(*[range(4)] * 4)))
    visit = (x.sum(axis=1) & 1).reshape(dims)
    assert pass_conflicts(nd, dims, visit) > 0


def interacts(p, q):
    """passes p, q read or write each other's links (parities 1 or 2 apart)"""
    return 1 <= bin(p ^ q).count("1") <= 2


@pytest.mark.parametrize("nd,nlaunch", [(2, 4), (3, 7), (4, 11)])
def test_paired_launch_order_keeps_every_dependency(nd, nlaunch):
    """The library's launch order for a sweep with independent passes two to
    a launch (klft_sweep_launches, the schedule the sweep kernels run): every
    pass once, and of two passes that touch each other's links the earlier
    one in the reference's order 0..2^nd-1 (Metropolis.hpp:139-162) runs in
    an earlier launch.  11 launches in 4D is the longest such chain
    (0 1 2 3 5 6 7 11 13 14 15)."""
    import klft_amd as klft
    launches = klft.sweep_launches(nd)
    assert len(launches) == nlaunch
    flat = [p for launch in launches for p in launch]
    assert sorted(flat) == list(range(1 << nd))
    at = {p: i for i, launch in enumerate(launches) for p in launch}
    for q in range(1 << nd):
        for p in range(q + 1, 1 << nd):
            if interacts(p, q):
                assert at[q] < at[p], (q, p)
    chain = [0, 1, 2, 3, 5, 6, 7, 11, 13, 14, 15]
    if nd == 4:
        assert all(interacts(a, b) for a, b in zip(chain, chain[1:]))


@pytest.mark.parametrize("nd,dims", [(3, (4, 6, 8)), (4, (4, 4, 6, 4)), (4, (8, 8, 8, 8)), (4, (2, 4, 2, 6))])
def test_paired_passes_are_race_free(nd, dims):
    """Each launch that runs two passes, emulated as one pass over both
    passes' sites: no link read on behalf of one site is written by another
    (extents of 2 included, where x + mu and x - mu coincide)."""
    import klft_amd as klft
    visit = schedule(nd, dims)
    merged = visit.copy()
    for launch in klft.sweep_launches(nd):
        if len(launch) == 2:
            merged[visit == launch[1]] = launch[0]
    assert pass_conflicts(nd, dims, merged) == 0


@pytest.mark.parametrize("pair", [(0, 1), (1, 2), (5, 6)])
def test_pairing_dependent_passes_would_race(pair):
    """The emulation sees the conflict when two passes one or two parities
    apart share a launch."""
    nd, dims = 4, (4, 4, 6, 4)
    visit = schedule(nd, dims)
    merged = visit.copy()
    merged[visit == pair[1]] = pair[0]
    assert pass_conflicts(nd, dims, merged) > 0


@pytest.mark.parametrize("nd,nsweeps,nlaunch", [(2, 5, 20), (3, 5, 31), (4, 1, 11), (4, 5, 43), (4, 50, 403)])
def test_run_launch_order_keeps_every_dependency(nd, nsweeps, nlaunch):
    """The library's launch order for a run of sweeps (klft_run_launches, the
    schedule metropolis_run follows): every (sweep, pass) once, at most two
    to a launch, the two of a launch three or more parities apart, and of two
    passes that touch the same links (parities at most two apart, the same
    sublattice of another sweep included) the earlier one in the reference's
    order (sweep by sweep, passes 0..2^nd-1, Metropolis.hpp:139-162) runs in
    an earlier launch.  One sweep is the single-sweep order; over many sweeps
    4D takes 8 launches per sweep instead of 11."""
    import klft_amd as klft
    launches = klft.run_launches(nd, nsweeps)
    assert len(launches) == nlaunch
    if nsweeps == 1:
        assert [tuple(p for p, _ in launch) for launch in launches] == klft.sweep_launches(nd)
    flat = [sp for launch in launches for sp in launch]
    assert sorted(flat) == sorted((p, s) for s in range(nsweeps) for p in range(1 << nd))
    at = {sp: i for i, launch in enumerate(launches) for sp in launch}
    for launch in launches:
        assert len(launch) <= 2
        if len(launch) == 2:
            assert bin(launch[0][0] ^ launch[1][0]).count("1") >= 3
    order = [(p, s) for s in range(nsweeps) for p in range(1 << nd)]
    for i, a in enumerate(order):
        for b in order[i + 1:]:
            if bin(a[0] ^ b[0]).count("1") <= 2:
                assert at[a] < at[b], (a, b)


@pytest.mark.parametrize("nd,dims", [(3, (4, 6, 8)), (4, (4, 4, 6, 4)), (4, (2, 4, 2, 6))])
def test_run_paired_passes_are_race_free(nd, dims):
    """Every pair of passes that a run of sweeps puts in one launch (passes of
    consecutive sweeps included), emulated as one pass over both passes'
    sites: no link read on behalf of one site is written by another."""
    import klft_amd as klft
    visit = schedule(nd, dims)
    pairs = {tuple(sorted(p for p, _ in launch)) for launch in klft.run_launches(nd, 4) if len(launch) == 2}
    assert pairs
    for a, b in sorted(pairs):
        merged = visit.copy()
        merged[visit == b] = a
        assert pass_conflicts(nd, dims, merged) == 0, (a, b)


def test_run_launches_reports_capacity_and_bad_arguments():
    """klft_run_launches fills at most `capacity` launches: too small a buffer
    is an error that still reports the count needed; nd outside 2..4 and a
    negative sweep count are errors; zero sweeps is zero launches."""
    import ctypes as C
    import klft_amd as klft
    lib = klft.load()
    n = C.c_int(-1)
    buf = (C.c_int32 * (4 * 19))()  # capacity counts launches, four ints each
    assert lib.klft_run_launches(4, 2, buf, 10, C.byref(n)) != 0
    assert n.value == 19
    assert lib.klft_run_launches(4, 2, buf, 19, C.byref(n)) == 0 and n.value == 19
    assert lib.klft_run_launches(5, 2, buf, 10, C.byref(n)) != 0
    assert lib.klft_run_launches(4, -1, buf, 10, C.byref(n)) != 0
    assert lib.klft_run_launches(4, 0, None, 0, C.byref(n)) == 0 and n.value == 0
    assert klft.run_launches(3, 0) == []


def greedy_run_schedule(nd, nsweeps):
    """The run schedule by its definition: passes in program order, each into
    the first launch after the last pass it conflicts with (parities at most
    two apart) that holds fewer than two passes."""
    launches, at = [], {}
    order = [(p, s) for s in range(nsweeps) for p in range(1 << nd)]
    for i, (p, s) in enumerate(order):
        lo = max([at[q] + 1 for q in order[:i] if bin(p ^ q[0]).count("1") <= 2], default=0)
        k = lo
        while k < len(launches) and len(launches[k]) >= 2:
            k += 1
        if k == len(launches):
            launches.append([])
        launches[k].append((p, s))
        at[(p, s)] = k
    return [tuple(launch) for launch in launches]


@pytest.mark.parametrize("nd", [2, 3, 4])
@pytest.mark.parametrize("nsweeps", [1, 2, 3, 7])
def test_run_launches_equal_the_greedy_schedule(nd, nsweeps):
    """The library walks the run schedule as a head, one window per further
    sweep and a tail (its period of one sweep); the walk equals the greedy
    schedule built pass by pass from its definition."""
    import klft_amd as klft
    assert klft.run_launches(nd, nsweeps) == greedy_run_schedule(nd, nsweeps)
