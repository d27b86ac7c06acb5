"""Batched right-hand sides of independent instances (uneven grid sizes and orbital counts) and
the contraction autotuner, against the FP64 oracle and against the instances' own launches."""
import numpy as np
import pytest

from codyfortran_amd import hip, workloads as wl
from tests.common import oracle_system, state_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances():
    import torch

    dev = torch.device("cuda:0")
    ws = wl.batched_workloads(n=6, seed=11, sizes=(512, 1024), orbitals=(4, 6, 8, 12))
    out = []
    for i, w in enumerate(ws):
        ctx, x, wts = wl.make_context(w)
        state = wl.seeded_state(x, wts, w.N, w.M, seed=100 + i)
        d_s = torch.from_numpy(state.view(np.float64).copy()).to(dev)
        d_o = torch.zeros_like(d_s)
        out.append((w, ctx, wts, state, d_s, d_o))
        print(f"\n[instance {w.name}: nG={w.nG} N={w.N} M={w.M} nC={w.nC}]", flush=True)
    yield out
    for _, ctx, *_ in out:
        ctx.close()


def test_batch_matches_instances_and_oracle(instances):
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    refs = []
    for w, ctx, wts, state, d_s, d_o in instances:
        ctx.set_graphs(False)
        ctx.time_derivative_dev(d_s.data_ptr(), d_o.data_ptr(), 0.0, stream)
        torch.cuda.synchronize()
        refs.append(d_o.cpu().numpy().copy())
        ctx.set_graphs(True)
    b = hip.Batch([inst[1] for inst in instances])
    try:
        for rep in range(3):
            for *_, d_o in instances:
                d_o.zero_()
            b.time_derivative_dev([i[4].data_ptr() for i in instances], [i[5].data_ptr() for i in instances],
                                  0.0, stream)
            torch.cuda.synchronize()
            for (w, ctx, wts, state, d_s, d_o), ref in zip(instances, refs):
                assert np.array_equal(d_o.cpu().numpy(), ref), (rep, w.name)
    finally:
        b.close()
    for (w, ctx, wts, state, d_s, d_o), ref in zip(instances, refs):
        s, _, _ = oracle_system(w)
        err = state_rel(ref.view(np.complex128), s.time_derivative(state), w.nC, wts, w.M)
        print(f"  {w.name}: RHS rel err {err:.3e}")
        assert err < 1e-12, w.name


@pytest.mark.parametrize("forms", ["int8", "mixed"])
def test_grouped_batch_bit_identical_to_instances(instances, forms):
    """The batch groups instances by orbital count and contraction path (FP64, tiled int8) into one
    launch per kernel type; instances on other paths (int8 at M = 12, untiled) keep their own
    launches.  Every instance's right-hand side is bit-identical to its own launches, with all
    instances on the int8-sliced contraction and with FP64 and int8 instances mixed."""
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    for k, (w, ctx, *_) in enumerate(instances):
        int8 = forms == "int8" or k % 2 == 0
        ctx.set_precision(hip.PREC_INT8_SLICED if int8 else hip.PREC_FP64)
    try:
        refs = []
        for w, ctx, wts, state, d_s, d_o in instances:
            ctx.set_graphs(False)
            ctx.time_derivative_dev(d_s.data_ptr(), d_o.data_ptr(), 0.0, stream)
            torch.cuda.synchronize()
            refs.append(d_o.cpu().numpy().copy())
            ctx.set_graphs(True)
        b = hip.Batch([inst[1] for inst in instances])
        try:
            for rep in range(3):
                for *_, d_o in instances:
                    d_o.zero_()
                b.time_derivative_dev([i[4].data_ptr() for i in instances],
                                      [i[5].data_ptr() for i in instances], 0.0, stream)
                torch.cuda.synchronize()
                for (w, ctx, wts, state, d_s, d_o), ref in zip(instances, refs):
                    assert np.array_equal(d_o.cpu().numpy(), ref), (forms, rep, w.name)
        finally:
            b.close()
        for (w, ctx, wts, state, d_s, d_o), ref in zip(instances, refs):
            s, _, _ = oracle_system(w)
            err = state_rel(ref.view(np.complex128), s.time_derivative(state), w.nC, wts, w.M)
            assert err < 1e-10, (forms, w.name, err)
    finally:
        for w, ctx, *_ in instances:
            ctx.set_precision(hip.PREC_FP64)


def test_autotune_keeps_parity(instances):
    w, ctx, wts, state, d_s, d_o = instances[-1]
    ctx.upload_state(state)
    k = ctx.autotune()
    assert k in (1, 2, 4, 8)
    s, _, _ = oracle_system(w)
    err = state_rel(ctx.time_derivative(state), s.time_derivative(state), w.nC, wts, w.M)
    print(f"  {w.name}: k split {k}, RHS rel err {err:.3e}")
    assert err < 1e-12


def test_full_batch_matches_oracle():
    """The 64 seeded instances of the batched workload (N_g up to 4096, M up to 12), each context
    autotuned, replayed as one graph: every instance within 1e-12 of the FP64 oracle.  (Also the
    regression test of the allocation fill that a kernel on a non-blocking stream could overtake,
    which left a context's tiled interaction matrix partly zeroed at random.)"""
    import torch

    dev = torch.device("cuda:0")
    ws = wl.batched_workloads(n=64)
    ctxs, states, d_s, d_o = [], [], [], []
    try:
        for i, w in enumerate(ws):
            ctx, x, wts = wl.make_context(w)
            st = wl.seeded_state(x, wts, w.N, w.M, seed=500 + i)
            ctx.upload_state(st)
            ctx.autotune()
            ctxs.append((ctx, wts))
            states.append(st)
            d_s.append(torch.from_numpy(st.view(np.float64).copy()).to(dev))
            d_o.append(torch.zeros_like(d_s[-1]))
        b = hip.Batch([c for c, _ in ctxs])
        stream = torch.cuda.current_stream().cuda_stream
        ps, po = [t.data_ptr() for t in d_s], [t.data_ptr() for t in d_o]
        for _ in range(2):
            b.time_derivative_dev(ps, po, 0.0, stream)
        torch.cuda.synchronize()
        first = [o.cpu().numpy().copy() for o in d_o]
        # the schedule autotuner keeps the results bit for bit (the grouped launches run the
        # instances' own bodies, whichever chain or group they land in)
        pick, ms = b.autotune(ps, po)
        print(f"  batch autotune: schedule {hip.Batch.SCHEDULES[pick]}, ms per batched RHS by schedule "
              + ", ".join(f"{m:.3f}" for m in ms))
        assert 0 <= pick < len(ms) and all(m > 0 for m in ms) and ms[pick] == min(ms)
        # the autotuned schedule, then fixed ones: one chain, and four with every group split
        for sched in (None, (1, 1.0), (4, 0.25)):
            if sched:
                b.set_schedule(*sched)
            for o in d_o:
                o.zero_()
            b.time_derivative_dev(ps, po, 0.0, stream)
            torch.cuda.synchronize()
            for w, o, f in zip(ws, d_o, first):
                assert np.array_equal(o.cpu().numpy(), f), (sched, w.name)
        with pytest.raises(hip.CodyError):
            b.set_schedule(5, 1.0)  # (four chains at most)
        b.close()
        worst = 0.0
        for (ctx, wts), w, st, o in zip(ctxs, ws, states, d_o):
            s, _, _ = oracle_system(w)
            err = state_rel(o.cpu().numpy().view(np.complex128), s.time_derivative(st), w.nC, wts, w.M)
            del s
            worst = max(worst, err)
            assert err < 1e-12, (w.name, err)
        print(f"  64 instances: worst RHS rel err {worst:.3e}")
    finally:
        for c, _ in ctxs:
            c.close()
