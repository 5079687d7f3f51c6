"""GPU: the endpoint sort entry point that allocates nothing
(fyzz_hip_sap_sort_endpoints_fixed, workspace in the library's fixed
per-device array) against the workspace-argument entry point
(fyzz_hip_sap_sort_endpoints), bit for bit, on two calls in a row (the
second smaller, so stale workspace contents from the first would show),
plus the oracle's sort of the same bounds; and the size limit rejected
before any launch."""
import ctypes as C

import numpy as np
import pytest

from fyzz_amd import scenes

pytestmark = pytest.mark.gpu


def _bounds(orc, torch, seed, n):
    half, pos = scenes.sap_scene(n, seed=seed, extent=2.0 * n ** (1 / 3))
    quat = scenes.normalize_quat32(scenes.random_quat(scenes.uniform(seed, 20, n), scenes.uniform(seed, 21, n),
                                                      scenes.uniform(seed, 22, n)))
    off = np.full(n, 0.02, np.float32)
    pos[n // 2:n // 2 + 7] = pos[0]  # equal keys: ties broken by slot
    half[n // 2:n // 2 + 7] = half[0]
    quat[n // 2:n // 2 + 7] = quat[0]
    want = orc.box_bounds(pos, quat, half, off)
    return want, torch.from_numpy(want.view(np.int32)).cuda()


def test_fixed_workspace_sort_matches_workspace_argument_sort(ctx, orc, torch_cuda):
    from fyzz_amd import sap
    from fyzz_amd._lib import lib
    torch = torch_cuda
    max_boxes = int(lib().fyzz_hip_sap_sort_fixed_max_boxes())
    assert max_boxes == 131072
    for seed, n in ((31, max_boxes), (32, 9_000)):
        want_b, b = _bounds(orc, torch, seed, n)
        fv, fp = sap.sort_endpoints_fixed(ctx, b)
        av, ap = sap.sort_endpoints(ctx, b)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(fv.cpu().numpy(), av.cpu().numpy(), err_msg=f"values n={n}")
        np.testing.assert_array_equal(fp.cpu().numpy(), ap.cpu().numpy(), err_msg=f"payloads n={n}")
        wv, wp = orc.sort_endpoints(want_b)
        np.testing.assert_array_equal(fv.cpu().numpy().view(np.uint32), wv, err_msg=f"oracle values n={n}")
        np.testing.assert_array_equal(fp.cpu().numpy().view(np.uint32), wp, err_msg=f"oracle payloads n={n}")


def test_fixed_workspace_rejects_oversize(ctx, torch_cuda):
    from fyzz_amd._lib import lib
    n = int(lib().fyzz_hip_sap_sort_fixed_max_boxes()) + 1
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lib().fyzz_hip_sap_sort_endpoints_fixed(ctx.h, C.c_uint32(n), dummy, dummy, dummy)
    assert rc == -1
    assert b"fixed workspace" in lib().fyzz_hip_last_error(ctx.h)
