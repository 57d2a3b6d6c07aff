"""GPU: pp_divergence_noalloc (device arrays, the reduction's partial sums in
the library's static scratch array) against pp_divergence (host arrays,
allocating) on two calls in a row, the second on a smaller grid (another
context on the same GPU, sharing the scratch array): nlu0 bit for bit and
the volume sum bit for bit (the same kernels and reduction order)."""
import numpy as np
import pytest

import cases
from flow_structure_interaction_fem_ibm_amd import pp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def one_call(n, ny, nz, seed, dev):
    print(f"[case] pp_divergence_noalloc {n}x{ny}x{nz} seed {seed}", flush=True)
    xn, yn, zn, ib, gs, ex = cases.sphere_case(n, r=0.2, stretch=1.3, ny=ny, nz=nz)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    rng = np.random.default_rng(seed)
    f = [rng.standard_normal(s.shape) for _ in range(3)]
    ref, ref_sum = s.divergence(*f, dtinv=2.5)
    tf = [torch.from_numpy(a).to(dev) for a in f]
    out = torch.full(s.shape, 7.0, dtype=torch.float64, device=dev)   # stale values must be overwritten
    got_sum = s.divergence_noalloc(*tf, out, dtinv=2.5)
    assert np.array_equal(out.cpu().numpy(), ref)
    assert got_sum == ref_sum
    # enqueue only (no sum): the same field once the stream has drained
    out2 = torch.zeros(s.shape, dtype=torch.float64, device=dev)
    assert s.divergence_noalloc(*tf, out2, dtinv=2.5, want_sum=False) is None
    torch.cuda.synchronize()
    assert np.array_equal(out2.cpu().numpy(), ref)
    return s


def test_divergence_noalloc_matches_the_allocating_entry():
    dev = torch.device("cuda:0")
    s1 = one_call(48, 40, 36, 1, dev)
    s2 = one_call(24, 20, 18, 2, dev)      # smaller: fewer partial sums in the same scratch array
    del s1, s2
