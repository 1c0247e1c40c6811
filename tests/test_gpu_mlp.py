"""Proposal network kernels against the CPU reference (GPU)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print_case(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("sizes,batch", [((64, 64, 32, 64), 32), ((32768, 512, 512, 512, 64), 32),
                                         ((32768, 512, 512, 512, 64), 256)])
def test_mlp_forward_backward_match_reference(sizes, batch):
    """MFMA forward/backward equal the numpy restatement that rounds the
    same operands to bf16 (only the accumulation order differs), and stay
    within bf16 error of plain FP32.  Batch 256 runs the layer-0 products on
    the macro-tiled kernel (forward split over K, weight gradient whole)."""
    from cuda_ndt_matcher_amd.proposal import ProposalNet
    from oracle import mlp_ref as R

    _print_case(f"mlp {sizes} batch {batch}")
    net = ProposalNet(sizes, seed=3)
    p = net.get_params()
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, (batch, sizes[0])).astype(np.float32) * (rng.uniform(0, 1, (batch, sizes[0])) < 0.1)
    dout = rng.normal(0, 1, (batch, sizes[-1])).astype(np.float32)
    out, g = net.forward_backward(x, dout)
    out_m, g_m = R.forward_backward(p, sizes, x, dout, mimic_bf16=True)
    out_f, g_f = R.forward_backward(p, sizes, x, dout, mimic_bf16=False)
    np.testing.assert_allclose(out, out_m, rtol=0, atol=2e-3 * np.abs(out_m).max())
    np.testing.assert_allclose(g, g_m, rtol=0, atol=2e-2 * np.abs(g_m).max())
    np.testing.assert_allclose(out, out_f, rtol=0, atol=3e-2 * np.abs(out_f).max())
    # the bulk agrees much more tightly than the worst entry
    assert np.median(np.abs(g - g_m)) <= 1e-3 * np.abs(g_m).max()


def test_bev_matches_reference():
    from cuda_ndt_matcher_amd import proposal
    from oracle import mlp_ref as R

    _print_case("bev")
    rng = np.random.default_rng(11)
    scans = [rng.uniform([-40, -40, -3], [40, 40, 16], (n, 3)).astype(np.float32) for n in (1000, 5000, 64)]
    got = proposal.bev_batch(scans).cpu().numpy().view(np.uint16)
    for j, s in enumerate(scans):
        ref = R.bf16(R.bev(s)).view(np.uint32) >> 16
        assert np.array_equal(got[j], ref.astype(np.uint16))
    # rows picked through an index, repeated, each rotated by its own heading
    store = proposal.ScanStore()
    store.add(scans)
    store.finish()
    rows = np.array([2, 0, 1, 0], np.int64)
    yaw = np.array([0.3, -2.0, math.pi / 2, 1.0])
    got = store.bev(rows, yaw).cpu().numpy().view(np.uint16)
    for j, (q, t) in enumerate(zip(rows, yaw)):
        rot = (np.float32(np.cos(t)), np.float32(np.sin(t)))
        ref = R.bf16(R.bev(scans[q], rot=rot)).view(np.uint32) >> 16
        assert np.array_equal(got[j], ref.astype(np.uint16)), (j, q, t)


def _min_error(out, labels):
    o = out.reshape(out.shape[0], -1, 4).astype(np.float64)
    e = (o[..., 0] - labels[:, :1]) ** 2 + (o[..., 1] - labels[:, 1:2]) ** 2 + 2 * (1 - np.cos(o[..., 2] - labels[:, 2:3]))
    return e.min(1).mean()


def test_training_reduces_the_proposal_loss():
    """Adam on the winner-takes-all loss fits a small synthetic mapping from
    inputs to poses: the error of the best hypothesis drops below a third of
    its start, and the loss the kernel reports equals the reference loss of
    the network's outputs."""
    import torch

    from cuda_ndt_matcher_amd.proposal import ProposalNet
    from oracle import mlp_ref as R

    _print_case("mlp training")
    sizes = (256, 128, 128, 64)
    net = ProposalNet(sizes, seed=5)
    rng = np.random.default_rng(13)
    B = 64
    x = rng.uniform(0, 1, (B, sizes[0])).astype(np.float32)
    A = rng.normal(0, 0.05, (sizes[0], 3))
    labels = np.tanh(x @ A).astype(np.float32)
    xb = torch.from_numpy(R.bf16(x).view(np.int32) >> 16).to(torch.int16).to("cuda:0")
    torch.cuda.synchronize()
    before = _min_error(net.forward_device(xb.data_ptr(), B), labels)
    for _ in range(300):
        net.train_step_device(xb.data_ptr(), labels, 3e-3)
    out = net.forward_device(xb.data_ptr(), B)
    after = _min_error(out, labels)
    assert after < 0.35 * before, (before, after)
    ref, _ = R.wta_loss(out, labels)
    assert abs(ref - net.train_step_device(xb.data_ptr(), labels, 0.0)) <= 1e-3 * max(1.0, abs(ref))


def test_query_scan_search_matches_the_batch_path():
    """QueryScan: the descriptor row equals bev_batch's, the padding rows stay
    zero, and a learned search through the device copy of the scan returns the
    same alignment as through the host scan."""
    import torch

    from cuda_ndt_matcher_amd import NdtScanMatcher, proposal, synth

    _print_case("query scan")
    wl = synth.make_workload("small", 0)
    m = NdtScanMatcher.builder().resolution(2.0).max_iterations(30).transformation_epsilon(0.01).build()
    m.set_target(wl.map_points)
    samples = proposal.make_samples(0, 5, 2, 40.0)
    net = proposal.ProposalNet(proposal.NET_SIZES, seed=1)
    frame = proposal.Frame(0.0, 0.0, 4.0)
    q = proposal.QueryScan(capacity=64)  # grows on the first scan
    for smp in samples:
        q.load(smp.scan)
        want = proposal.bev_batch([smp.scan])
        torch.cuda.synchronize()
        assert torch.equal(q.batch[:1], want)
        assert not torch.any(q.batch[1:])
        a, fa = proposal.learned_search(m, net, frame, smp, None, query=q)
        b, fb = proposal.learned_search(m, net, frame, smp, want)
        assert fa == fb and np.array_equal(a.pose_vector, b.pose_vector) and a.iterations == b.iterations
