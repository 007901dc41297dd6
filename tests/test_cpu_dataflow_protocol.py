"""The dataflow sweep's claim protocol (sweep.hip, sweep_dataflow), simulated
on the CPU over a real factor's forward dependency structure -- as the
kernel runs it (no claim window) and with a claim window, a design candidate
for bounding the number of polling waves.

Batches are up to 64 consecutive rows of one topological-depth level, in
sweep order, as the device batch table holds them (a batch never straddles a
level, so it depends only on batches of earlier levels). Resident waves
claim `chunk` consecutive batches per ticket from one counter and run them in
order; with a claim window W a ticket starts only while it lies fewer than W
batches past the count of finished batches (added per finished ticket). In
each simulated step every wave acts once, in a shuffled order: it claims,
waits on the window, or polls its current batch's dependencies and
publishes the batch when all of them were published in an earlier step.

Checked: every configuration finishes (no wave waits forever on a ticket
nobody can finish), every batch is published exactly once and after all of
its dependencies; without a window the polling population grows with the
number of resident waves (every claimed ticket polls), and a window bounds
the tickets whose waves poll row slots to ceil(W / chunk) + 1. A windowed
run (512 waves, a window of 8 tickets) finishes within a few steps of the
unwindowed one with less than half its polls."""
import numpy as np
import pytest

import apxchol_amd as ax
import oracle
from apxchol_amd import matgen as mg


def batch_dag():
    A = mg.add_excess(mg.grid_laplacian((20, 20, 20), "lognormal"), 1e-3)
    F = ax.factorize(A, seed=0)
    o = oracle.Oracle(A, F.export(), F.sddm, use_rounds=False)
    a = o.arrays()
    ptr, idx = a["csr_ptr"].astype(np.int64), a["csr_idx"].astype(np.int64)
    m = o.m
    depth = np.zeros(m, np.int64)
    deps = []
    for i in range(m):
        d = idx[ptr[i]:ptr[i + 1] - 1]   # diagonal last
        deps.append(d)
        depth[i] = depth[d].max() + 1 if d.size else 0
    order = np.lexsort((np.arange(m), depth))          # level by level, rows ascending
    pos = np.empty(m, np.int64)
    pos[order] = np.arange(m)
    batches, batch_of = [], np.empty(m, np.int64)
    i = 0
    while i < m:
        j = i
        while j < m and j - i < 64 and depth[order[j]] == depth[order[i]]:
            j += 1
        batch_of[order[i:j]] = len(batches)
        batches.append(order[i:j])
        i = j
    bdeps = [sorted({int(batch_of[d]) for r in rows for d in deps[r]}) for rows in batches]
    return bdeps


def simulate(bdeps, waves, chunk, window, seed=0):
    nb = len(bdeps)
    rng = np.random.default_rng(seed)
    published = np.full(nb, -1, np.int64)   # step of publication
    ticket = 0
    finished = 0
    state = [None] * waves                  # (b0, b1, next batch) of the wave's ticket
    polls = 0
    max_polling = 0
    step = 0
    while finished < nb:
        step += 1
        assert step < 50 * nb, "no progress"
        polling = 0
        done_now = []
        for w in rng.permutation(waves):
            if state[w] is None:
                if ticket >= nb:
                    continue
                state[w] = [ticket, min(ticket + chunk, nb), ticket]
                ticket += chunk
            b0, b1, b = state[w]
            if window and b0 - finished >= window:
                continue
            polling += 1
            polls += 1
            if all(0 <= published[d] < step for d in bdeps[b]):
                assert published[b] == -1
                published[b] = step
                state[w][2] = b + 1
                if b + 1 == b1:
                    done_now.append(b1 - b0)
                    state[w] = None
        finished += sum(done_now)
        max_polling = max(max_polling, polling)
    assert np.all(published > 0)
    for b, ds in enumerate(bdeps):
        assert all(published[d] < published[b] for d in ds)
    return step, polls, max_polling


@pytest.fixture(scope="module")
def dag():
    return batch_dag()


def test_dependencies_point_backwards(dag):
    assert all(d < b for b, ds in enumerate(dag) for d in ds)
    assert len(dag) > 100


@pytest.mark.parametrize("waves", [1, 7, 64, 512])
@pytest.mark.parametrize("chunk", [1, 2, 8])
def test_claim_protocol_finishes_and_window_bounds_the_pollers(dag, waves, chunk):
    for window in (0, chunk, 64):
        steps, polls, max_polling = simulate(dag, waves, chunk, window)
        if window:
            assert max_polling <= -(-window // chunk) + 1, (window, max_polling)


def test_window_cuts_polls_not_progress(dag):
    steps0, polls0, _ = simulate(dag, 512, 2, 0)
    steps1, polls1, _ = simulate(dag, 512, 2, 16)
    assert steps1 <= steps0 + 8
    assert polls1 * 2 < polls0


def test_without_a_window_every_claimed_ticket_polls(dag):
    """The kernel's protocol: the peak number of polling waves follows the
    resident grid (all of them poll from the moment they claim)."""
    peaks = [simulate(dag, w, 2, 0)[2] for w in (8, 64, 512)]
    assert peaks[0] <= 8 and peaks[1] == 64 and peaks[2] >= 80
