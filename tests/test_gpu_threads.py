"""Several factor objects driven from several host threads at once.

Kept in a file of its own, after the parity suite: it is the one test whose
outcome depends on the host's thread schedule, and a failure here must not
hide the parity tests.
"""
import numpy as np
import pytest

from hifde3d_amd import problems as pb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import hifde3d_amd
    hifde3d_amd.lib()
    return hifde3d_amd


def test_factor_objects_from_several_threads_match_sequential(H):
    """Factor objects driven from several host threads at once (each on its
    own stream, graph captures included) give the same bits as one after
    another.  Each worker produces and consumes its vectors on a PyTorch
    stream of its own and synchronizes only that stream: a device-wide
    synchronize would be refused by HIP while another thread captures a
    graph (the library's own calls never issue one)."""
    import threading
    import torch
    probs = [(32, 4, s) for s in range(4)] + [(24, 3, 7)]
    facs, B = [], []
    for NC, P, seed in probs:
        f = H.HIFFactor(NC - 1, P, 1e-4, precision="mixed")
        f.set_option("host_threads", 2)
        f.set_stencil(pb.pow10_coeff(NC, seed, -2.0, 2.0), pb.zeros_coeff(NC), True)
        facs.append(f)
        B.append(torch.from_numpy(pb.rhs_normal(NC, 50 + seed, 4)).cuda())
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in facs]

    def run(i, out):
        f = facs[i]
        with torch.cuda.stream(streams[i]):
            assert f.Factor() == 0
            x = torch.zeros_like(B[i])
            it, rr = f.pcg_multi_device(B[i].data_ptr(), x.data_ptr(), 4, tol=1e-10, maxit=200)
            xh = x.cpu()
            streams[i].synchronize()
        out[i] = (list(it), xh.numpy())

    seq = {}
    for i in range(len(facs)):
        run(i, seq)
    for rep in range(2):
        par, errs = {}, []

        def work(i):
            try:
                run(i, par)
            except Exception as e:  # reported below with the worker's index
                errs.append((i, repr(e)))
        th = [threading.Thread(target=work, args=(i,)) for i in range(len(facs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, (rep, errs)
        assert sorted(par) == list(range(len(facs))), rep
        for i in range(len(facs)):
            assert par[i][0] == seq[i][0], (rep, i)
            assert np.array_equal(par[i][1], seq[i][1]), (rep, i)
    torch.cuda.synchronize()
    for f in facs:
        f.close()
