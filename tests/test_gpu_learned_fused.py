"""The learned search in one launch (adock_mc_learned: chain set up once, the proposal network run on MFMA by each
chain's wavefront inside its step loop) against the step-by-step pipeline (one adock_mlp_infer + one adock_mc_step
launch per MC step) and against the fused plain kernel: bit for bit in poses and energies, the same logged moves, the
same features, and a search split in two calls equal to one call."""
import numpy as np
import pytest
import torch

from adock_amd import experiments, synth, workload
from adock_amd.kernel import KernelProblem
from adock_amd.learned import StepSearch
from adock_amd.mlp import ProposalNet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_cx():
    w, prob = workload.single("small", seed=0)
    kp = KernelProblem(prob)
    N = 32
    t, q, th = synth.start_states(4, N, w["center"], w["size"], w["lig"].n_tors, kp.KM)
    st = kp.pack(torch.as_tensor(t).cuda(), torch.as_tensor(q).cuda(), torch.as_tensor(th).cuda())
    return w, kp, st


@pytest.fixture(scope="module")
def proposal_cx():
    """Four complexes of the learned-proposal workload (4,000-atom receptor, 35-atom ligand each), 64 chains each."""
    _, ligs, prob = experiments.complexes(range(20_000, 20_004))
    kp = KernelProblem(prob)
    st, cl = workload.chain_starts(3, ligs, [np.zeros(3)] * len(ligs), [np.full(3, 26.0)] * len(ligs), 64, kp.KM)
    return kp, torch.as_tensor(st).cuda(), torch.as_tensor(cl).cuda()


def _bufs(st):
    N = st.shape[0]
    return st.clone(), torch.zeros(N).cuda(), st.clone(), torch.full((N,), float("inf")).cuda()


def _assert_same(a_bufs, b_bufs):
    for a, b in zip(a_bufs, b_bufs):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def _net(seed, w3_scale):
    net = ProposalNet("cuda", seed=seed)
    net.p["W3"].mul_(w3_scale)
    net.p["b3"].copy_(torch.linspace(-0.3, 0.3, net.p["b3"].numel(), device="cuda"))
    net.refresh_bf16()
    return net


def test_fused_plain_equals_fused_mc_kernel(small_cx):
    print("[gpu] adock_mc_learned without the network vs adock_mc", flush=True)
    w, kp, st = small_cx
    N, S = st.shape[0], 15
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    ref = _bufs(st)
    kp.run_mc(cl, torch.arange(N, dtype=torch.int32).cuda(), *ref, True, S, 0, 1.2, 2.0, 2.0, 777, 0)
    got = _bufs(st)
    StepSearch(kp, cl).fused(*got, S, seed=777)
    torch.cuda.synchronize()
    _assert_same(ref, got)


@pytest.mark.parametrize("w3_scale,p_plain", [(1.0, 0.25), (20.0, 0.0)])
def test_fused_network_search_equals_step_pipeline(small_cx, w3_scale, p_plain):
    print(f"[gpu] fused learned search vs step pipeline, W3 x {w3_scale}, p_plain {p_plain}", flush=True)
    w, kp, st = small_cx
    N, S = st.shape[0], 10
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    net = _net(1, w3_scale)
    step = StepSearch(kp, cl, net=net)
    a = _bufs(st)
    step.run(*a, S, seed=5, use_net=True, p_plain=p_plain)
    fused = StepSearch(kp, cl, net=net)
    b = _bufs(st)
    fused.fused(*b, S, seed=5, use_net=True, p_plain=p_plain)
    torch.cuda.synchronize()
    _assert_same(a, b)
    np.testing.assert_array_equal(step.feat.cpu().numpy(), fused.feat.cpu().numpy())
    # the network moved the chains: the search differs from the plain one
    c = _bufs(st)
    StepSearch(kp, cl).fused(*c, S, seed=5)
    assert not torch.equal(b[0], c[0])


def test_fused_log_holds_the_step_pipelines_records(small_cx):
    print("[gpu] fused move log vs step-pipeline log", flush=True)
    w, kp, st = small_cx
    N, S = st.shape[0], 12
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    step = StepSearch(kp, cl, log_cap=20000)
    a = _bufs(st)
    step.run(*a, S, seed=9, log=True)
    fused = StepSearch(kp, cl, log_cap=20000)
    b = _bufs(st)
    fused.fused(*b, S, seed=9, log=True)
    torch.cuda.synchronize()
    _assert_same(a, b)
    ra = torch.cat(step.logged()[:3] + (step.logged()[3][:, None],), 1).cpu().numpy()
    rb = torch.cat(fused.logged()[:3] + (fused.logged()[3][:, None],), 1).cpu().numpy()
    assert ra.shape == rb.shape and ra.shape[0] > 50
    sa = ra[np.lexsort(ra.T[::-1])]
    sb = rb[np.lexsort(rb.T[::-1])]
    np.testing.assert_array_equal(sa, sb)


def test_fused_split_equals_one_call_and_invariants(small_cx):
    print("[gpu] fused learned search: two calls vs one", flush=True)
    w, kp, st = small_cx
    N = st.shape[0]
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    net = _net(2, 5.0)
    one = _bufs(st)
    StepSearch(kp, cl, net=net).fused(*one, 12, seed=21, use_net=True, p_plain=0.25)
    two = _bufs(st)
    ss = StepSearch(kp, cl, net=net)
    ss.fused(*two, 5, seed=21, use_net=True, p_plain=0.25)
    ss.fused(*two, 7, seed=21, step_offset=5, init_opt=False, use_net=True, p_plain=0.25)
    torch.cuda.synchronize()
    _assert_same(one, two)
    e_best, _ = kp.energy_grad(cl, one[2])
    e_cur, _ = kp.energy_grad(cl, one[0])
    torch.testing.assert_close(e_best, one[3], rtol=1e-4, atol=1e-4)
    torch.testing.assert_close(e_cur, one[1], rtol=1e-4, atol=1e-4)
    assert (one[3] <= one[1] + 1e-6).all()


def test_fused_equals_step_pipeline_on_the_proposal_workload(proposal_cx):
    print("[gpu] fused vs step pipeline, 4 proposal-workload complexes x 64 chains", flush=True)
    kp, st, cl = proposal_cx
    net = _net(3, 5.0)
    a = _bufs(st)
    StepSearch(kp, cl, net=net).run(*a, 6, seed=13, use_net=True, p_plain=0.25)
    b = _bufs(st)
    StepSearch(kp, cl, net=net).fused(*b, 6, seed=13, use_net=True, p_plain=0.25)
    torch.cuda.synchronize()
    _assert_same(a, b)


def _bias_net(scores):
    """A network whose output is its last-layer bias alone (W3 = 0): output j = scores.get(j, 0)."""
    net = ProposalNet("cuda", seed=0)
    net.p["W3"].zero_()
    b = torch.zeros_like(net.p["b3"])
    for j, v in scores.items():
        b[j] = v
    net.p["b3"].copy_(b)
    net.refresh_bf16()
    return net


@pytest.mark.parametrize("out,entity", [(0, 0), (3, 1), (8, 3)])
def test_entity_choice_follows_the_scores(small_cx, out, entity):
    """net_mode 1 with one entity scored far above the others and a small tau: every network move perturbs that
    entity (output 0 -> translation, 3 -> rotation, 7 + k -> torsion k), and every move is logged with log_all."""
    print(f"[gpu] entity choice: output {out} -> entity {entity}", flush=True)
    w, kp, st = small_cx
    assert w["lig"].n_tors >= 2
    N, S = st.shape[0], 6
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    ss = StepSearch(kp, cl, net=_bias_net({out: 10.0}), log_cap=N * S + 64)
    ss.fused(*_bufs(st), S, seed=17, use_net=True, p_plain=0.0, net_mode=1, tau=0.1, log=True, log_all=True)
    torch.cuda.synchronize()
    X, T, M, gain, ent, bgap = ss.logged(entity=True)
    assert (bgap.cpu().numpy() <= 0).all()
    assert len(ent) == N * S
    assert (ent.cpu().numpy() == entity).all()
    # with p_plain = 1 the entities are the plain move's uniform draw over 2 + nt
    ss.log_count.zero_()
    ss.fused(*_bufs(st), S, seed=17, use_net=True, p_plain=1.0, net_mode=1, tau=0.1, log=True, log_all=True)
    e = ss.logged(entity=True)[4].cpu().numpy()
    assert len(e) == N * S and e.min() >= 0 and e.max() < 2 + w["lig"].n_tors and len(set(e.tolist())) > 2


def test_fused_entity_search_equals_step_pipeline(small_cx):
    print("[gpu] fused entity-choice search vs step pipeline, every move logged", flush=True)
    w, kp, st = small_cx
    N, S = st.shape[0], 10
    cl = torch.zeros(N, dtype=torch.int32).cuda()
    net = _net(5, 5.0)
    step = StepSearch(kp, cl, net=net, log_cap=N * S + 64)
    a = _bufs(st)
    step.run(*a, S, seed=23, use_net=True, p_plain=0.25, net_mode=1, tau=0.2, log=True, log_all=True)
    fused = StepSearch(kp, cl, net=net, log_cap=N * S + 64)
    b = _bufs(st)
    fused.fused(*b, S, seed=23, use_net=True, p_plain=0.25, net_mode=1, tau=0.2, log=True, log_all=True)
    torch.cuda.synchronize()
    _assert_same(a, b)
    ra = step.log_buf[:int(step.log_count.item())].cpu().numpy()
    rb = fused.log_buf[:int(fused.log_count.item())].cpu().numpy()
    assert ra.shape == rb.shape == (N * S, ra.shape[1])
    np.testing.assert_array_equal(ra[np.lexsort(ra.T[::-1])], rb[np.lexsort(rb.T[::-1])])
