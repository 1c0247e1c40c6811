"""The map-independent learned proposal (prior_proposal.py) on the GPU, at a
reduced budget: trained here on two fresh procedural scenes, tested on a third
that it never saw.

What training gains: the hypotheses' heading error modulo quarter turns (what
a scan of these axis-aligned scenes shows of its heading) is at least halved
against the same network untrained.  Against the plain search: the learned
search (hypotheses, polish, and the plain 200-particle search with the
prior's roll/pitch as fallback) succeeds at least as often as the plain
200-particle search of the configuration (roll/pitch N(0, 1 deg)), each
within 2 cm / 0.1 deg of the alignment started at the true pose."""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print_case(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def test_prior_proposal_gains_on_an_unseen_scene():
    from cuda_ndt_matcher_amd import NdtScanMatcher, NdtScanMatcherConfig, prior_proposal as pp, proposal, synth

    _print_case("prior proposal: training set (2 scenes)")
    data = pp.TrainingSet()
    for i in range(2):
        data.add_scene(2000 + i, 900 + i, 2048, matcher_factory=lambda: NdtScanMatcher(NdtScanMatcherConfig()))
    data.finish()
    trained = proposal.ProposalNet(proposal.NET_SIZES, seed=3)
    untrained = proposal.ProposalNet(proposal.NET_SIZES, seed=3)
    _print_case("prior proposal: training")
    losses = pp.train_relative(trained, data, 40, seed=3)
    assert losses[-1] < losses[0]
    del data
    _print_case("prior proposal: unseen scene")
    scene = synth.Scene(2077, 200.0, 40, 100)
    reqs = pp.scene_requests(scene, 4242, 24, 60.0, 64, 900, pp.DESC_LEAF)
    m = NdtScanMatcher(NdtScanMatcherConfig())
    m.set_target(scene.sample(200_000, 0.02, 2077))
    refs = pp.reference_poses(m, reqs)
    e_trained = pp.heading_error_mod_quarter(trained, reqs, refs)
    e_untrained = pp.heading_error_mod_quarter(untrained, reqs, refs)
    _print_case(f"heading error mod 90 deg, p50: trained {np.median(e_trained):.2f}, "
                f"untrained {np.median(e_untrained):.2f}")
    assert np.median(e_trained) <= 0.5 * np.median(e_untrained)
    dev = pp.DeviceRequest()
    ok_plain = ok_learned = 0
    t_plain = t_learned = 0.0
    for k, (q, ref) in enumerate(zip(reqs, refs)):
        t = time.perf_counter()
        r = pp.plain_prior_search(m, q, 200, 50 + k)
        t_plain += time.perf_counter() - t
        ok_plain += pp.success(r.pose_vector, ref)
        t = time.perf_counter()
        r, _ = pp.relative_search(m, trained, q, dev,
                                  fallback=lambda q_, k=k: pp.plain_prior_search(m, q_, 200, 50 + k,
                                                                                 roll_pitch_deg=0.0))
        t_learned += time.perf_counter() - t
        ok_learned += pp.success(r.pose_vector, ref)
    _print_case(f"success: plain {ok_plain}/{len(reqs)} in {t_plain:.2f} s, learned {ok_learned}/{len(reqs)} "
                f"in {t_learned:.2f} s")
    m.close()
    trained.close()
    untrained.close()
    assert ok_learned >= ok_plain
