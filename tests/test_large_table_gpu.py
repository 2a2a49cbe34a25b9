"""An f32 SH3 table of 25,165,824 splats: more than the 23,860,929 splats per
plane that one 4 GiB window over all 45 SH planes could address (the
projection reads the SH planes through three resources of 15 planes). The
fused project + sort is held to SortKeys.sort of the keys it wrote
(src/core/sort.ts:64-163), bit-exact, and a sample of splats spread over
the table (half of it in the last million, whose SH planes lie more than
4 GiB past the first SH plane) to the oracle's projection of the same
splats (projection_check.compare: keys, word fields within one code,
floats within 1e-5)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector, workloads as W
from tests.projection_check import compare

pytestmark = pytest.mark.gpu

N_BIG = 25_165_824
SIDE_PLANES = {"conic": 3, "cov2d": 3, "opacity": 1, "ndc": 2, "view_z": 1, "bound": 3, "rgb": 3, "distance": 1}


def test_f32_table_beyond_one_plane_window(cuda):
    assert N_BIG * 4 * 45 > 0xFFFFFFFF
    spec = W.plain_4m_spec(seed=5, count=N_BIG)
    planes = W.generate(spec)
    rig = W.plain_4m_camera(seed=5)
    prm = Params(sort_bits=32)
    table = DeviceTable.from_f32_planes(planes, N_BIG, 3)
    print(f"f32 SH3 table of {N_BIG:,} splats, project_sort with side outputs", flush=True)
    out = Projector().project_sort(table, rig, prm, side=True)
    torch.cuda.synchronize()
    keys = out["keys"].cpu().numpy().view(np.uint32)[:N_BIG]
    count = int(out["count"].cpu().item())
    order = out["indices"].cpu().numpy().view(np.uint32)[:count]
    ordering, visible = O.sort(keys, 32)
    print(f"visible {visible:,}", flush=True)
    assert visible > N_BIG // 2
    assert count == visible
    assert np.array_equal(order, ordering[:visible])
    assert list(out["args"].cpu().numpy()) == [6144, -(-visible // 1024), 0, 0, 0]

    rng = np.random.default_rng(5)
    idx = np.unique(np.concatenate([rng.integers(0, N_BIG, 4096), rng.integers(N_BIG - 1_000_000, N_BIG, 4096)]))
    sub = np.ascontiguousarray(planes[:, idx])
    ref = O.project_f32(sub, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    side = {}
    for name, k in SIDE_PLANES.items():
        side[name] = out["side"][name].cpu().numpy().reshape(k, -1)[:, idx].reshape(-1)
    g = {"words": [out["words"][0].cpu().numpy().view(np.uint32)[idx]], "keys": keys[idx], "side": side}
    rep = compare(g, ref, 1)
    print(rep, flush=True)
    assert rep["eye0_compared"] > idx.size // 2
