"""The minimal H.264 decoder (libpcme_h264.so, include/pcme_dec.h) against
the reference's own fixture, and the oracle's motion compensation and MV
prediction against the decoder.

* tests/golden/h264/baseline.h264 is the reference's x264 stream
  (tests/refs/h264/manifest.toml:1-14: Baseline profile, CAVLC, P frames,
  2 slices per picture, 3 reference frames, keyint 5, 176x144, 10 frames;
  file sha256 da7318e9...).  Its decoded frames must hash to the sha256 of
  x264's own reconstruction recorded in the manifest (5a978004...ef00), as
  the reference's harness checks (src/h264/tests.rs:171-189).
* With the decoder's motion compensation and MV prediction pinned that way
  (the stream codes fractional MVs of 15 of the 16 quarter phases,
  16x16 / 16x8 / 8x16 / 8x8 partitions, several references and P_Skip),
  the oracle's
  o_hpel + o_mc_luma and o_mvp / o_pskip_mv are checked against them on the
  decoded pictures: reference-held vectors for both."""
from __future__ import annotations

import ctypes
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
STREAM = os.path.join(HERE, "golden", "h264", "baseline.h264")
STREAM_SHA256 = "da7318e9fb42977072b45b9cfc88a5f453d641ae7a3381dfacc7eb0617248c87"
RECON_SHA256 = "5a978004cf205415f059349e9fa14714e1d7578b1c5165c48520fca71a11ef00"


@pytest.fixture(scope="module")
def dec():
    from pcme import h264dec
    h264dec.load()
    return h264dec


@pytest.fixture(scope="module")
def stream():
    data = open(STREAM, "rb").read()
    assert hashlib.sha256(data).hexdigest() == STREAM_SHA256
    return data


@pytest.fixture(scope="module")
def frames(dec, stream):
    return dec.decode_annexb(stream)


def test_library_exports_every_declared_symbol(dec):
    import re
    text = open(os.path.join(HERE, "..", "include", "pcme_dec.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = {m.group(1) for m in re.finditer(r"\b(pcme_\w+)\s*\(", text)}
    assert names == set(dec.EXPORTS), names ^ set(dec.EXPORTS)
    lib = dec.load()
    for n in names:
        assert hasattr(lib, n), n


def test_decodes_bit_exact_with_the_x264_reconstruction(dec, frames):
    assert len(frames) == 10
    assert all((f.width, f.height) == (176, 144) for f in frames)
    assert dec.frames_digest(frames) == RECON_SHA256


def test_whole_stream_in_one_call_and_byte_chunks(dec, stream):
    """The same frames whether the stream comes as one buffer or as access
    units (NAL units are never split)."""
    assert dec.frames_digest(dec.decode_annexb(stream, per_access_unit=False)) == RECON_SHA256


def test_stream_exercises_what_the_pins_below_rely_on(frames):
    """Two slices, I and P pictures, every P partition mode x264 was
    allowed, several reference indices, P_Skip and fractional MVs of 15 of
    the 16 quarter phases (not (3, 2), the sample 'k' = (j + m + 1) >> 1,
    which the spec check of the QPEL_SRC table in test_abi.py covers), so
    the MC and MVP checks below cover those cases."""
    from pcme.h264dec import KIND_P, KIND_PSKIP
    mbs = np.concatenate([f.mbs for f in frames])
    assert set(np.unique(mbs["slice"])) == {0, 1}
    p = mbs[mbs["kind"] == KIND_P]
    assert set(np.unique(p["part"])) >= {0, 1, 2, 3}
    assert (p["ref"] > 0).any() and (mbs["kind"] == KIND_PSKIP).any()
    phases = set(map(tuple, (p["mv"].reshape(-1, 2) & 3).tolist()))
    assert len(phases) == 15 and (3, 2) not in phases


def test_rejects_what_is_outside_the_subset(dec, stream):
    """A PPS with entropy_coding_mode_flag set (the fixture's PPS with that
    one bit flipped) is refused with a message, as are garbage and a
    picture cut short."""
    from pcme import PcmeError
    i = stream.index(b"\x00\x00\x00\x01\x68") + 5
    cabac = bytearray(stream)
    assert cabac[i] == 0xCB  # pps_id 0, sps_id 0, entropy_coding_mode_flag 0, ...
    cabac[i] |= 0x20
    with pytest.raises(PcmeError, match="CABAC"):
        dec.decode_annexb(bytes(cabac))
    cut = stream[:stream.index(b"\x00\x00\x00\x01\x41", 1500) + 40]
    with pytest.raises(PcmeError):
        dec.decode_annexb(cut)
    d = dec.Decoder()
    assert d.decode(b"") == [] and d.flush() == []
    d.close()


def test_oracle_mc_luma_matches_the_decoder(dec, frames, oracle):
    """o_hpel + o_mc_luma (the encoder's padded planes and QPEL_SRC
    averages) equal the decoder's 8.4.2.2.1 interpolation on the decoded
    pictures for blocks of every search size at every quarter phase,
    inside the picture and up to 40 pixels beyond its edges."""
    rs = np.random.default_rng(5)
    lib = oracle.load()
    mbw, mbh = 11, 9
    checked = 0
    for f in frames[:4]:
        ref = oracle.Ref(oracle.pad_luma(f.y, mbw, mbh))
        for _ in range(150):
            bw, bh = [(16, 16), (16, 8), (8, 16), (8, 8)][rs.integers(0, 4)]
            x, y = int(rs.integers(0, 176 - bw + 1)), int(rs.integers(0, 144 - bh + 1))
            mvx = int(rs.integers(-4 * (x + 40), 4 * (176 - bw - x + 40) + 1))
            mvy = int(rs.integers(-4 * (y + 40), 4 * (144 - bh - y + 40) + 1))
            got = np.zeros((bh, bw), np.uint8)
            lib.o_mc_luma(ctypes.byref(ref.view), x, y, oracle.omv(mvx, mvy), bw, bh, oracle._p(got), bw)
            want = dec.mc_luma(f.y, x, y, (mvx, mvy), bw, bh)
            assert np.array_equal(got, want), (x, y, mvx, mvy, bw, bh)
            checked += 1
    assert checked == 600


def test_oracle_mv_prediction_matches_the_decoder(frames, oracle):
    """o_mvp (16x16, 16x8, 8x16 and 8x8 partitions, directional rules,
    neighbours within the MB already searched) and o_pskip_mv equal the
    predictors the decoder derived for x264's MBs, with the decoded MB
    field as the neighbourhood (intra neighbours, several reference
    indices, skipped MBs).  MBs with a neighbour in the other slice are
    left out (the encoder codes one slice per picture)."""
    from pcme.h264dec import KIND_I4x4, KIND_I16x16, KIND_IPCM, KIND_P, KIND_PSKIP
    lib = oracle.load()
    mbw, mbh = 11, 9
    prm = oracle.oparams(mbw, mbh, 16, 1, 1, 16 * mbh + 44)
    kind_map = {KIND_I4x4: 1, KIND_I16x16: 3, KIND_IPCM: 3, KIND_P: 5, KIND_PSKIP: 4}
    n_part = n_skip = 0
    for f in frames:
        mbs = f.mbs
        if not (mbs["kind"] >= KIND_P).any():
            continue
        field = (oracle.ombinfo * (mbw * mbh))()
        for i, m in enumerate(mbs):
            field[i].kind = kind_map[int(m["kind"])]
            for k in range(4):
                field[i].ref_idx[k] = int(m["ref"][k])
            for k in range(16):
                field[i].mv[k] = oracle.omv(int(m["mv"][k][0]), int(m["mv"][k][1]))
        for my in range(mbh):
            for mx in range(mbw):
                m = mbs[my * mbw + mx]
                nbs = [(mx - 1, my), (mx, my - 1), (mx + 1, my - 1), (mx - 1, my - 1)]
                if any(0 <= x < mbw and 0 <= y < mbh and mbs[y * mbw + x]["slice"] != m["slice"] for x, y in nbs):
                    continue
                if m["kind"] == KIND_PSKIP:
                    got = lib.o_pskip_mv(ctypes.byref(prm), field, mx, my)
                    assert (got.x, got.y) == tuple(m["mvp"][0]), (mx, my)
                    n_skip += 1
                    continue
                if m["kind"] != KIND_P or (m["part"] == 3 and (m["sub"] != 0).any()):
                    continue
                cur = oracle.ombinfo()
                cur.kind = 5
                for k in range(4):
                    cur.ref_idx[k] = int(m["ref"][k])
                part = int(m["part"])
                # (x, y, w) in 4x4 units, shape code, 4x4 blocks of each partition
                parts = {0: [(0, 0, 4, 0)], 1: [(0, 0, 4, 1), (0, 2, 4, 2)], 2: [(0, 0, 2, 3), (2, 0, 2, 4)],
                         3: [(0, 0, 2, 0), (2, 0, 2, 0), (0, 2, 2, 0), (2, 2, 2, 0)]}[part]
                wh = {0: (4, 4), 1: (4, 2), 2: (2, 4), 3: (2, 2)}[part]
                mask = 0
                for x, y, w, shape in parts:
                    blk = y * 4 + x
                    rf = int(m["ref"][(y // 2) * 2 + x // 2])
                    got = lib.o_mvp(ctypes.byref(prm), field, ctypes.byref(cur), mask, mx, my, x, y, w, rf, shape)
                    assert (got.x, got.y) == tuple(m["mvp"][blk]), (mx, my, part, x, y)
                    n_part += 1
                    for yy in range(y, y + wh[1]):
                        for xx in range(x, x + wh[0]):
                            cur.mv[yy * 4 + xx] = oracle.omv(int(m["mv"][yy * 4 + xx][0]), int(m["mv"][yy * 4 + xx][1]))
                            mask |= 1 << (yy * 4 + xx)
    assert n_part > 200 and n_skip > 20, (n_part, n_skip)
