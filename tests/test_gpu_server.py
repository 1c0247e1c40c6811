"""Continuous-batching server: results alone and in the mix (GPU).

A hipFree anywhere in the process waits for every launch on the device to
end, the server's persistent launch included.  The tests therefore close
their matchers explicitly and collect garbage before starting a server, so
that no matcher left over from an earlier test is destroyed (and its buffers
freed) while a server runs."""
import gc
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = 8  # workgroups per slot, and per standalone alignment for the comparison


def _print_case(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def small():
    from cuda_ndt_matcher_amd import synth

    return synth.make_workload("small", 0)


def _requests(small):
    rng = np.random.default_rng(41)
    n_all = small.scan.shape[0]
    reqs = []
    for k, n in enumerate((300, 1400, 700, 1200, 90, 1000, 450, 1400)):
        n = min(n, n_all)
        sel = np.sort(rng.choice(n_all, n, replace=False))
        pose = small.guess + np.array([0.05 * (k % 3), -0.04 * (k % 2), 0.0, 0.0, 0.0, 0.01 * (k % 4)])
        reqs.append((np.ascontiguousarray(small.scan[sel]), pose))
    return reqs


def _drain(srv, got, want, deadline):
    while len(got) < want and time.time() < deadline:
        for t, r in srv.poll():
            got[t] = r
    return got


def test_server_results_match_standalone(small):
    """Eight uneven requests through three slots, four submitted at first and
    four while those run: every result equals the standalone alignment of the
    same scan and start with the same number of workgroups."""
    from cuda_ndt_matcher_amd import NdtScanMatcher, NdtScanMatcherConfig

    _print_case("server vs standalone")
    m = NdtScanMatcher(NdtScanMatcherConfig(workgroups_per_alignment=W))
    m.set_target(small.map_points)
    reqs = _requests(small)
    alone = [m.align(s, p) for s, p in reqs]
    got = {}
    gc.collect()
    deadline = time.time() + 60
    with m.server(slots=3, workgroups_per_slot=W, max_points=2000) as srv:
        tickets = [srv.submit(s, p) for s, p in reqs[:4]]  # one more than the slots: it waits in the queue
        _drain(srv, got, 2, deadline)
        tickets += [srv.submit(s, p) for s, p in reqs[4:]]  # admitted while earlier ones run
        _drain(srv, got, len(reqs), deadline)
    m.close()
    assert len(got) == len(reqs), "server did not finish every request"
    for t, a in zip(tickets, alone):
        r = got[t]
        assert r.status == a.status and r.iterations == a.iterations
        np.testing.assert_allclose(r.pose_vector, a.pose_vector, rtol=0, atol=1e-9)
        assert r.nvtl == pytest.approx(a.nvtl, rel=1e-12, abs=1e-12)
        assert r.num_correspondences == a.num_correspondences


def test_server_errors(small):
    from cuda_ndt_matcher_amd import NdtError, NdtScanMatcher, NdtScanMatcherConfig

    _print_case("server errors")
    m = NdtScanMatcher(NdtScanMatcherConfig())
    m.set_target(small.map_points)
    with pytest.raises(NdtError) as e:
        m.server(slots=4096, workgroups_per_slot=64, max_points=100)
    assert e.value.code == -7
    gc.collect()
    with m.server(slots=2, workgroups_per_slot=4, max_points=100) as srv:
        with pytest.raises(NdtError) as e:
            srv.submit(small.scan[:101], small.guess)
        assert e.value.code == -7
        with pytest.raises(NdtError) as e:
            srv.submit(small.scan[:0], small.guess)
        assert e.value.code == -2
        t = srv.submit(small.scan[:100], small.guess)
        got = _drain(srv, {}, 1, time.time() + 30)
        assert t in got
    m.close()


def test_handle_refuses_other_work_while_serving(small):
    """A running server holds the device's co-resident workgroups: the
    handle's other calls and a second server are refused with an error (not
    left waiting), and everything works again after stop."""
    from cuda_ndt_matcher_amd import NdtError, NdtScanMatcher, NdtScanMatcherConfig

    _print_case("server guard")
    m = NdtScanMatcher(NdtScanMatcherConfig(workgroups_per_alignment=W))
    m.set_target(small.map_points)
    want = m.align(small.scan, small.guess)
    gc.collect()
    srv = m.server(slots=2, workgroups_per_slot=W, max_points=small.scan.shape[0])
    try:
        with pytest.raises(NdtError):
            m.align(small.scan, small.guess)
        with pytest.raises(NdtError):
            m.evaluate_nvtl(small.scan, small.guess)
        with pytest.raises(NdtError):
            m.set_target(small.map_points)
        with pytest.raises(NdtError):
            m.server(slots=1, workgroups_per_slot=W, max_points=100)
        t = srv.submit(small.scan, small.guess)
        got = _drain(srv, {}, 1, time.time() + 60)
        assert t in got
    finally:
        srv.stop()
    r = m.align(small.scan, small.guess)
    m.close()
    assert np.array_equal(r.pose_vector, want.pose_vector) and got[t].iterations == want.iterations


def test_server_idle_timeout_is_launch_wide(small):
    """With a short idle timeout, a slot that never gets work does not leave
    while the host keeps calling (requests keep landing on slot 0 one at a
    time); a later burst still reaches both slots.  Once the host stops
    calling for longer than the timeout, the whole launch ends: submit and
    poll raise, stop succeeds, and the handle works again."""
    from cuda_ndt_matcher_amd import NdtError, NdtScanMatcher, NdtScanMatcherConfig

    _print_case("server idle timeout")
    m = NdtScanMatcher(NdtScanMatcherConfig(workgroups_per_alignment=W))
    m.set_target(small.map_points)
    scan = np.ascontiguousarray(small.scan[:800])
    want = m.align(scan, small.guess)
    gc.collect()
    srv = m.server(slots=2, workgroups_per_slot=W, max_points=scan.shape[0])
    try:
        # A timeout of a few seconds: the host's own pauses between calls
        # (scheduling, the interpreter) can reach a few hundred milliseconds.
        srv.set_idle_timeout(2.0)
        t_end = time.time() + 6.0  # three timeouts' worth of serving on one slot
        n_done = 0
        while time.time() < t_end:
            t = srv.submit(scan, small.guess)
            got = _drain(srv, {}, 1, time.time() + 30)
            assert t in got and got[t].iterations == want.iterations
            n_done += 1
        burst = [srv.submit(scan, small.guess) for _ in range(2)]
        got = _drain(srv, {}, 2, time.time() + 30)
        assert set(burst) <= set(got)
        time.sleep(3.0)  # no calls: the launch ends
        with pytest.raises(NdtError) as e:
            srv.submit(scan, small.guess)
        assert e.value.code == -5
    finally:
        srv.stop()
    assert n_done >= 3
    r = m.align(scan, small.guess)
    m.close()
    assert np.array_equal(r.pose_vector, want.pose_vector)
