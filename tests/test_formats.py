"""On-disk formats without ITK: NRRD (detached header) and MetaImage round trips,
and the NRRD header written for a scalar image against the reference's own
fixture header (Testing/Data/Input/Bullseyes/BullseyeTestBlur00.nhdr, copied under
tests/golden), line for line apart from the data file name."""
import os

import numpy as np

from atlaswerks_amd import metaimage

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_nrrd_scalar_roundtrip(tmp_path):
    a = np.random.default_rng(0).random((5, 6, 7)).astype(np.float32)
    p = str(tmp_path / "a.nhdr")
    metaimage.write_nrrd(p, metaimage.Image(a, (2.0, 2.0645161, 1.7777778), (-277.8, -198.6, 54.2)))
    b = metaimage.read_nrrd(p)
    assert np.array_equal(a, b.data)
    assert b.size == (7, 6, 5)
    np.testing.assert_allclose(b.spacing, (2.0, 2.0645161, 1.7777778), rtol=1e-7)
    np.testing.assert_allclose(b.origin, (-277.8, -198.6, 54.2), rtol=1e-7)


def test_nrrd_vector_roundtrip(tmp_path):
    v = np.random.default_rng(1).random((4, 5, 6, 3)).astype(np.float32)
    p = str(tmp_path / "h.nhdr")
    metaimage.write_nrrd(p, metaimage.Image(v, (1.0, 1.5, 2.0), (1.0, 2.0, 3.0)))
    w = metaimage.read_nrrd(p)
    assert w.data.shape == (4, 5, 6, 3)
    assert np.array_equal(v, w.data)
    assert w.spacing == (1.0, 1.5, 2.0)


def test_nrrd_header_matches_reference_fixture(tmp_path):
    src = os.path.join(GOLDEN, "BullseyeTestBlur00.nhdr")
    ref = metaimage.read_nrrd(src)
    p = str(tmp_path / "BullseyeTestBlur00.nhdr")
    metaimage.write_nrrd(p, ref)
    again = metaimage.read_nrrd(p)
    assert np.array_equal(ref.data, again.data)

    def fields(path):
        out = {}
        for line in open(path):
            line = line.strip()
            if ":" in line and not line.startswith("#"):
                k, val = line.split(":", 1)
                out[k.strip()] = val.strip()
        return out
    want, got = fields(src), fields(p)
    for k in ("type", "dimension", "space", "sizes", "space directions", "kinds", "endian",
              "encoding", "space origin", "data file"):
        assert got[k] == want[k], (k, got[k], want[k])


def test_metaimage_roundtrip(tmp_path):
    v = np.random.default_rng(2).random((3, 4, 5, 3)).astype(np.float32)
    p = str(tmp_path / "f.mha")
    metaimage.write_metaimage(p, metaimage.Image(v, (1.0, 2.0, 3.0), (0.5, -1.0, 2.0)))
    w = metaimage.read_metaimage(p)
    assert np.array_equal(v, w.data)
    assert w.spacing == (1.0, 2.0, 3.0) and w.origin == (0.5, -1.0, 2.0)
