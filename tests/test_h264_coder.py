"""The field coder (libpcme_h264.so, include/pcme_enc.h): every stream it
writes decodes, through the minimal decoder pinned by x264's own
reconstruction, to exactly the coder's reconstruction; and its rate and
distortion behave as a coder's must (QP, skips, the field's quality)."""
from __future__ import annotations

import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def codec():
    from pcme import h264dec, h264enc
    h264enc.load()
    return h264enc, h264dec


def _clip(pcme_lib, w, h, n, seed=0x5EED7700, vmax=3.0):
    return pcme_lib.synth_clip(pcme_lib.ClipSpec(w, h, n, nrect=4, rect_lo=24, rect_hi=80, vmax=vmax, seed=seed))


def _zero_field(pcme_lib, w, h, n):
    mbw, mbh = pcme_lib.mb_dims(w, h)
    return np.zeros((n - 1, mbh, mbw), pcme_lib.MB_DTYPE)


def _roundtrip(dec, coded):
    frames = dec.decode_annexb(coded.stream)
    assert len(frames) == len(coded.recon)
    for f, r in zip(frames, coded.recon):
        assert np.array_equal(f.y, r[0]) and np.array_equal(f.u, r[1]) and np.array_equal(f.v, r[2])
    return frames


def test_exports_match_the_header(codec):
    import re
    enc, _ = codec
    text = open(os.path.join(HERE, "..", "include", "pcme_enc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = {m.group(1) for m in re.finditer(r"\b(pcme_\w+)\s*\(", text)}
    assert names == set(enc.EXPORTS)


@pytest.mark.parametrize("qp", [22, 27, 37])
def test_searched_field_decodes_to_the_reconstruction(codec, pcme_lib, oracle, qp):
    """The oracle's ME-only field (all four partition modes) coded at QP
    22/27/37: the decoder's output equals the coder's reconstruction."""
    enc, dec = codec
    y, u, v = _clip(pcme_lib, 320, 240, 5)
    field = oracle.search_clip_me(y, 16, 1, True, qp)
    assert set(np.unique(field["part"])) == {0, 1, 2, 3}
    c = enc.code_clip(y, u, v, field, qp)
    frames = _roundtrip(dec, c)
    assert all((f.width, f.height) == (320, 240) for f in frames)
    assert 30.0 < c.p_psnr < 60.0 and c.p_bits > 0


def test_cropped_size_and_all_modes_of_a_zero_field(codec, pcme_lib):
    """200x120 (coded 208x128, cropped in the SPS) with an all-zero field."""
    enc, dec = codec
    y, u, v = _clip(pcme_lib, 200, 120, 3)
    c = enc.code_clip(y, u, v, _zero_field(pcme_lib, 200, 120, 3), 30)
    frames = _roundtrip(dec, c)
    assert all((f.width, f.height) == (200, 120) for f in frames)
    assert np.array_equal(frames[0].y, y[0])  # the IDR is I_PCM: lossless


def test_rate_and_distortion_follow_qp(codec, pcme_lib, oracle):
    enc, dec = codec
    y, u, v = _clip(pcme_lib, 176, 144, 4)
    field = oracle.search_clip_me(y, 16, 1, True, 27)
    pts = [enc.code_clip(y, u, v, field, qp) for qp in (20, 26, 32, 38)]
    bits = [p.p_bits for p in pts]
    psnr = [p.p_psnr for p in pts]
    assert bits == sorted(bits, reverse=True) and psnr == sorted(psnr, reverse=True)


def test_a_static_clip_is_all_skips(codec, pcme_lib):
    """Every picture equal to the first and a zero field: each P MB's skip MV
    is zero and nothing survives quantisation, so every P MB is P_Skip and a
    P picture is a few bytes."""
    enc, dec = codec
    y, u, v = _clip(pcme_lib, 176, 144, 1)
    y, u, v = (np.repeat(a, 4, axis=0) for a in (y, u, v))
    e = enc.Encoder(176, 144)
    e.picture(y[0], u[0], v[0], None, 27)
    f = _zero_field(pcme_lib, 176, 144, 2)[0]
    for t in range(1, 4):
        b = e.picture(y[t], u[t], v[t], f, 27)
        assert e.stats() == {"skip": 99, "inter": 0, "pcm": 0} and len(b) < 16
    e.close()


def test_a_searched_field_costs_fewer_bits_than_no_motion(codec, pcme_lib, oracle):
    """On moving content the search's field codes in fewer bits than a zero
    field at the same QP (and no worse PSNR than 0.3 dB)."""
    enc, _ = codec
    y, u, v = _clip(pcme_lib, 320, 240, 4, vmax=6.0, seed=0x5EED7701)
    searched = enc.code_clip(y, u, v, oracle.search_clip_me(y, 16, 1, True, 27), 27)
    zero = enc.code_clip(y, u, v, _zero_field(pcme_lib, 320, 240, 4), 27)
    assert searched.p_bits < 0.9 * zero.p_bits and searched.p_psnr > zero.p_psnr - 0.3


def test_rejects_bad_input(codec, pcme_lib):
    from pcme import PcmeError
    enc, _ = codec
    y, u, v = _clip(pcme_lib, 64, 48, 1)
    e = enc.Encoder(64, 48)
    f = _zero_field(pcme_lib, 64, 48, 2)[0]
    with pytest.raises(PcmeError, match="before the first IDR"):
        e.picture(y[0], u[0], v[0], f, 27)
    e.picture(y[0], u[0], v[0], None, 27)
    f["mv"][0, 0, 3] = (4096, 0)
    with pytest.raises(PcmeError, match="2048"):
        e.picture(y[0], u[0], v[0], f, 27)
    f["mv"][0, 0, 3] = (0, 0)
    f["part"][0, 1] = 1
    f["mv"][0, 1, 9] = (4, 0)  # lower 16x8 partition: its blocks must share one MV
    with pytest.raises(PcmeError, match="differ inside partition"):
        e.picture(y[0], u[0], v[0], f, 27)
    with pytest.raises(PcmeError):
        enc.Encoder(63, 48)
    e.close()


@pytest.mark.parametrize("nref,qp", [(2, 27), (4, 32)])
def test_multiref_field_decodes_to_the_reconstruction(codec, pcme_lib, oracle, nref, qp):
    """The oracle's multi-reference field (o_search_frame_mr) coded with
    nref reference indices: the decoder's output equals the coder's
    reconstruction, and each coded P MB carries the field's reference."""
    from pcme.h264dec import KIND_P
    enc, dec = codec
    y, u, v = _clip(pcme_lib, 176, 144, 6, vmax=4.0)
    y = y.copy()
    y[2:] = ((y[2:].astype(int) + y[:-2].astype(int)) // 2).astype(np.uint8)
    field = oracle.search_clip_mr(y, nref, 16, 1, True, qp)
    assert (field["ref"] > 0).any()
    c = enc.code_clip(y, u, v, field, qp, nref=nref)
    frames = _roundtrip(dec, c)
    coded = 0
    for t, f in enumerate(frames[1:]):
        mbs = f.mbs.reshape(field[t].shape)
        p = mbs["kind"] == KIND_P
        assert np.array_equal(mbs["ref"][p][:, 0], field[t]["ref"][p])
        coded += int(p.sum())
    assert coded > 50


def test_older_reference_saves_bits(codec, pcme_lib, oracle):
    """Pictures that alternate between two scenes: with four references each
    P picture from the third on is coded from the identical picture two
    back, at a fraction of the one-reference bits."""
    enc, dec = codec
    a, b = (_clip(pcme_lib, 176, 144, 1, seed=s) for s in (0x5EED7710, 0x5EED7711))
    y, u, v = (np.concatenate([a[k], b[k], a[k], b[k], a[k]]) for k in range(3))
    one = enc.code_clip(y, u, v, oracle.search_clip_mr(y, 1, 16, 1, True, 27), 27)
    four = enc.code_clip(y, u, v, oracle.search_clip_mr(y, 4, 16, 1, True, 27), 27, nref=4)
    _roundtrip(dec, one)
    _roundtrip(dec, four)
    assert four.p_bits < 0.6 * one.p_bits and four.p_psnr >= one.p_psnr - 0.1


def test_multiref_rejects_an_index_beyond_those_in_use(codec, pcme_lib):
    from pcme import PcmeError
    enc, _ = codec
    y, u, v = _clip(pcme_lib, 64, 48, 3)
    e = enc.Encoder(64, 48, nref=2)
    e.picture(y[0], u[0], v[0], None, 27)
    f = _zero_field(pcme_lib, 64, 48, 2)[0]
    f["ref"][0, 0] = 1  # one reconstruction so far: only index 0 is in use
    with pytest.raises(PcmeError, match="beyond"):
        e.picture(y[1], u[1], v[1], f, 27)
    f["ref"][0, 0] = 0
    e.picture(y[1], u[1], v[1], f, 27)
    f["ref"][0, 0] = 1
    e.picture(y[2], u[2], v[2], f, 27)  # two reconstructions now
    e.close()
