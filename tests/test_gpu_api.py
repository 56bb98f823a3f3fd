"""The C ABI's state and transfer contract on the GPU: a rejected grid leaves
the context unusable rather than inconsistent, stage timings are never lost
to a small buffer, the asynchronous map download equals the synchronous one,
host and device particle uploads give identical fields, and the work
counters count what the stages did.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx():
    from avism_amd._lib import Context
    return Context(0)


def _particles(n, L0, seed):
    from avism_amd import synth
    return synth.zeldovich(n, L0, seed=seed)


def test_rejected_grid_leaves_no_stale_state():
    from avism_amd._lib import AvismError
    p = _particles(16, 32.0, 3)
    c = _ctx()
    c.set_grid(16, 32.0, 1)
    c.load_particles(*p.numpy())
    c.particles_to_grid(k=32)
    c.find_voids(0)
    c.merge_and_post()
    print("[rejected grid] 1200^3 periodic (ext 1500^3 cells)", flush=True)
    with pytest.raises(AvismError) as e:
        c.set_grid(1200, 32.0, 1)  # 1500^3 extended cells: above the 2^31 limit
    assert e.value.code == -4
    for call in (lambda: c.find_voids(0), lambda: c.marca(), lambda: c.fields(), lambda: c.merge_and_post()):
        with pytest.raises(AvismError) as e:
            call()
        assert e.value.code == -5, e.value  # AVISM_ERR_STATE
    # a valid grid makes the context usable again
    c.set_grid(16, 32.0, 1)
    c.load_particles(*p.numpy())
    c.particles_to_grid(k=32)
    assert c.find_voids(0) >= 0


def test_timings_beyond_the_first_buffer():
    """More stage records than the binding's first buffer holds: nothing is
    dropped, and the library reports the size it needs."""
    import ctypes as C
    from avism_amd._lib import lib
    p = _particles(16, 32.0, 5)
    c = _ctx()
    c.set_grid(16, 32.0, 1)
    c.load_particles(*p.numpy())
    c.set_timing(True)
    for _ in range(40):  # 7 records per field stage
        c.particles_to_grid(k=32)
    n = C.c_int32(4)
    ms = np.zeros(4, np.float32)
    names = C.create_string_buffer(64)
    rc = lib().avism_get_timings(c._h, names, len(names), ms.ctypes.data_as(C.c_void_p), C.byref(n))
    assert rc == -4 and n.value > 256, (rc, n.value)
    want = n.value
    tim = c.timings()
    assert len(tim) == want
    assert sum(1 for s, _ in tim if s == "vvel_sph_vw") == 40
    assert all(ms_ >= 0.0 for _, ms_ in tim)
    assert c.timings() == []  # consumed


@pytest.mark.parametrize("periodic", [0, 1])
def test_host_upload_and_async_maps_match(periodic):
    import torch
    from avism_amd._lib import AVISM_DEVICE
    n, L0 = 24, 48.0
    p = _particles(n, L0, 11)
    arr = p.numpy()
    c1 = _ctx()
    c1.set_grid(n, L0, periodic)
    c1.load_particles(*arr)  # pageable host arrays
    c1.particles_to_grid(k=32)
    f1 = c1.fields()
    _, h1 = c1.particle_fields(arr[0].size)
    nv1 = c1.find_voids(0)
    c1.merge_and_post()
    m1 = c1.marca(inner_only=True)
    # device arrays, then pinned host arrays with the asynchronous download
    dev = [torch.from_numpy(a).cuda() for a in arr]
    pin = [torch.from_numpy(a).pin_memory() for a in arr]
    for src, where in ((dev, AVISM_DEVICE), (pin, None)):
        c2 = _ctx()
        c2.set_grid(n, L0, periodic)
        c2.load_particles(*src, where=where)
        c2.particles_to_grid(k=32)
        u1 = torch.empty(n ** 3, dtype=torch.float32).pin_memory()
        dv = torch.empty(n ** 3, dtype=torch.float32).pin_memory()
        c2.fields_async(u1, dv)
        nv2 = c2.find_voids(0)
        c2.merge_and_post()
        m2 = torch.empty(n ** 3, dtype=torch.int32).pin_memory()
        c2.marca(inner_only=True, out=m2)
        c2.sync()
        _, h2 = c2.particle_fields(arr[0].size)
        assert np.array_equal(u1.numpy(), f1["u1co"])
        assert np.array_equal(dv.numpy(), f1["div"])
        assert np.array_equal(h2, h1)
        assert nv2 == nv1
        assert np.array_equal(m2.numpy(), m1)
        # a second field stage waits for the pending downloads before it writes
        c2.fields_async(u1, dv)
        c2.particles_to_grid(k=32)
        c2.sync()
        assert np.array_equal(u1.numpy(), f1["u1co"])


@pytest.mark.parametrize("periodic", [0, 1])
def test_inner_marca_is_the_ext_box(periodic):
    """avism_get_marca(inner_only) is the n^3 block of the ext map that starts
    at cell (1,1,1), read into host memory and into device memory."""
    import ctypes as C
    import torch
    from avism_amd._lib import AVISM_DEVICE, lib
    n, L0 = 24, 48.0
    p = _particles(n, L0, 21 + periodic)
    c = _ctx()
    c.set_grid(n, L0, periodic)
    c.load_particles(*p.numpy())
    c.particles_to_grid(k=32)
    nv = c.find_voids(0)
    c.merge_and_post()
    gi = c.grid()
    print(f"[inner marca] periodic={periodic} n={n} ext={gi.next} cubes={nv}", flush=True)
    ext = c.marca().reshape(gi.next, gi.next, gi.next)
    o = 1 - gi.low1
    want = ext[o:o + n, o:o + n, o:o + n].reshape(-1)
    assert (want > 0).any()
    assert np.array_equal(c.marca(inner_only=True), want)
    d = torch.full((n ** 3,), -7, dtype=torch.int32, device="cuda")
    assert lib().avism_get_marca(c._h, C.c_void_p(d.data_ptr()), 1, AVISM_DEVICE) == 0
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), want)


def test_counters():
    n, L0 = 20, 40.0
    p = _particles(n, L0, 13)
    arr = p.numpy()
    c = _ctx()
    c.set_grid(n, L0, 1)
    c.load_particles(*arr)
    c.particles_to_grid(k=32)
    cnt = c.counters()
    assert cnt["ppart_queries"] == arr[0].size
    assert cnt["vvel_queries"] == n ** 3
    # every particle's support holds at least its home cell's neighbourhood; at
    # one particle per cell and k = 32 the supports span tens of cells
    assert 8 * arr[0].size < cnt["ddens_pairs"] < 400 * arr[0].size
    mask = np.ones(n ** 3, np.int8)
    mask[: n ** 3 // 4] = 0
    c.set_mask(mask)
    c.particles_to_grid(k=32)
    assert c.counters()["vvel_queries"] == int(mask.sum())
