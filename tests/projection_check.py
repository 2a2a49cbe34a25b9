"""Comparison of a GPU projection with the oracle's, to the parity bar:

- sort keys: bit-exact, except where a cull decision sits within 1e-5
  (relative) of its threshold in the oracle (the visible flag may flip);
- words: each quantized field within one code (the reference allows a code
  "one step off" between f32 and f64 arithmetic, docs/encoding.md:88); the
  bound angle only where the ellipse is not near-circular (an angle of a
  circle is arbitrary);
- floats: conic within 1e-5 of its largest component, compensated opacity,
  view depth, ndc and key distance within 1e-5 relative, colours within 1e-5
  of max(|c|, 1) (colours are clamped at 0, so relative error near 0 is
  meaningless; 1 is the colour scale).
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

TOL = 1e-5


def _fields(words):
    w = words.astype(np.int64)
    return {
        "ndc_x": w[:, 0] & 0xFFFF, "ndc_y": w[:, 0] >> 16,
        "depth": w[:, 1] & 0xFFFFF, "angle": w[:, 1] >> 20,
        "opacity": w[:, 3] & 0xFF, "s1": (w[:, 3] >> 8) & 0xFFF, "s2": w[:, 3] >> 20,
    }


def _decode_rgb(words):
    L = O.lib()
    out = np.zeros((words.shape[0], 3))
    buf = np.zeros(3)
    for i, e in enumerate(words[:, 2]):
        L.orc_decode_ext_rgb(int(e), O.dptr(buf))
        out[i] = buf
    return out


def compare(gpu: dict, ref: dict, eyes: int, per_eye: bool = False, check_rgb_words: bool = True,
            rgb_sample: int = 20000, exact_rgb: bool = False) -> dict:
    """Assert parity; returns counts of what was compared and excused.

    exact_rgb: the colours are the stored base colours (no SH), so the
    encoder sees the same floats and the rgb words must be bit-exact."""
    report = {}
    margin = ref["margin"]
    near = margin < TOL
    gk = gpu["keys"]
    rk = ref["keys"]
    if per_eye:
        bad = np.nonzero((gk != rk).any(axis=0))[0]
    else:
        bad = np.nonzero(gk != rk)[0]
    unexcused = bad[~near[bad]]
    assert unexcused.size == 0, f"{unexcused.size} keys differ away from a cull threshold, first {unexcused[:5]}"
    report["keys_excused"] = int(bad.size)
    for e in range(eyes):
        gw = gpu["words"][e]
        rw = ref["words"][e]
        gv = (gw[:, 3] & 0xFF) != 0
        rv = ref["visible"][e].astype(bool)
        flip = np.nonzero(gv != rv)[0]
        assert np.all(near[flip]), f"eye {e}: {np.count_nonzero(~near[flip])} visibility flips away from thresholds"
        both = np.nonzero(gv & rv)[0]
        gf, rf = _fields(gw[both]), _fields(rw[both])
        for k in ("ndc_x", "ndc_y", "depth", "opacity", "s1", "s2"):
            d = np.abs(gf[k] - rf[k])
            assert d.max(initial=0) <= 1, f"eye {e} word field {k}: off by {d.max()} codes"
        # angle: only for clearly elongated ellipses (eigenvalue gap above 1e-3)
        da = np.abs(gf["angle"] - rf["angle"])
        da = np.minimum(da, 4096 - da)
        cs = gpu["side"]["bound"].reshape(eyes, 3, -1)[e]
        sc1, sc2 = cs[0][both], cs[1][both]
        elongated = (sc1 - sc2) > 1e-3 * sc1 * 2
        assert da[elongated].max(initial=0) <= 1, f"eye {e}: angle off by {da[elongated].max()} codes"
        report[f"eye{e}_compared"] = int(both.size)
        report[f"eye{e}_flips"] = int(flip.size)
        if check_rgb_words and both.size:
            sel = both[:: max(1, both.size // rgb_sample)]
            if exact_rgb:
                assert np.array_equal(gw[sel, 2], rw[sel, 2]), f"eye {e}: rgb words differ"
            gr, rr = _decode_rgb(gw[sel]), _decode_rgb(rw[sel])
            # one step of the coarser of the two codes: 2^(base - 15) / 255
            base = np.maximum(gw[sel, 2] >> 27, rw[sel, 2] >> 27).astype(np.float64)
            step = np.exp2(base - 15)[:, None] / 255
            assert np.all(np.abs(gr - rr) <= step * 1.0001 + 1e-12), f"eye {e}: rgb word beyond one step"
        # floats
        side = gpu["side"]
        n = gw.shape[0]
        conic = side["conic"].reshape(eyes, 3, -1)[e][:, both].astype(np.float64)
        rconic = ref["conic"][e][:, both]
        scale = np.abs(rconic).max(axis=0)
        assert np.all(np.abs(conic - rconic) <= TOL * scale), f"eye {e}: conic beyond 1e-5"
        op = side["opacity"].reshape(eyes, -1)[e][both].astype(np.float64)
        assert np.all(np.abs(op - ref["opacity"][e][both]) <= TOL * ref["opacity"][e][both]), "opacity"
        vz = side["view_z"].reshape(eyes, -1)[e][both].astype(np.float64)
        assert np.all(np.abs(vz - ref["view_z"][e][both]) <= TOL * np.abs(ref["view_z"][e][both])), "view z"
        ndc = side["ndc"].reshape(eyes, 2, -1)[e][:, both].astype(np.float64)
        rndc = ref["ndc"][e][:, both]
        assert np.all(np.abs(ndc - rndc) <= TOL * np.abs(rndc) + 1e-12), "ndc"
    n = gpu["words"][0].shape[0]
    vis_any = np.any(ref["visible"].astype(bool), axis=0)
    rgb = gpu["side"]["rgb"].reshape(3, -1)[:, vis_any].astype(np.float64)
    rr = ref["rgb"][:, vis_any]
    assert np.all(np.abs(rgb - rr) <= TOL * np.maximum(np.abs(rr), 1.0)), "rgb"
    dist = np.asarray(gpu["side"]["distance"], dtype=np.float64).reshape(-1)[: ref["distance"].size]
    dist = dist.reshape(ref["distance"].shape)
    assert np.all(np.abs(dist - ref["distance"]) <= TOL * ref["distance"]), "distance"
    return report


def to_host(out: dict, n: int, eyes: int) -> dict:
    """GPU projection outputs to numpy (uint32 words and keys, float sides)."""
    h = {"words": [w[:n].cpu().numpy().view(np.uint32) for w in out["words"]]}
    k = out["keys"]
    h["keys"] = k.cpu().numpy().view(np.uint32)
    if "side" in out:
        h["side"] = {name: t.cpu().numpy() for name, t in out["side"].items()}
    return h
