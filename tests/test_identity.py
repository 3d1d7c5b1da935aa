"""Content identities of the dense stage's inputs (reconstruction_stages.rs:458-510,
scene_project.rs:463-475), pinned by the mathematics: SHA-256 equals Python's hashlib on the
same bytes (and the FIPS 180-4 examples), and the request hash equals hashlib over the byte
layout request_content_hash_with_control feeds it, which the test builds independently with
struct. Also the reference's own test (tests/reconstruction_stages.rs:459-482): the identity
covers pixels, frame order, a single pixel and the options. Host only."""
import hashlib
import struct

import numpy as np
import pytest

from vt3d_dense import scenes

FIPS = [(b"abc", "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
        (b"", "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
        (b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq",
         "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1")]


def _frames(seed=3, n=4, size=(68, 48)):
    rng = np.random.default_rng(seed)
    W, H = size
    return [rng.integers(0, 256, (H, W, 4), dtype=np.uint8) for _ in range(n)]


def _expected(frames, o):
    """The bytes of request_content_hash_with_control (reconstruction_stages.rs:479-502)."""
    h = hashlib.sha256()
    h.update(b"video-to-3d/reconstruction-request/v1\0")
    h.update(struct.pack("<Q", o.max_features))
    h.update(struct.pack("<III", o.min_feature_distance, o.descriptor_radius, o.match_radius))
    h.update(struct.pack("<f", o.max_descriptor_distance))
    h.update(struct.pack("<f", o.ratio_threshold))
    h.update(b"\1" + struct.pack("<f", o.focal_length_pixels) if o.has_focal_length_pixels
             else b"\0")
    h.update(struct.pack("<QQ", o.max_dense_working_set_bytes, len(frames)))
    for f in frames:
        H, W = f.shape[:2]
        h.update(struct.pack("<IIQ", W, H, W * H * 4))
        h.update(np.ascontiguousarray(f).tobytes())
    return "sha256:" + h.hexdigest()


def test_sha256_known_answers(product):
    for msg, digest in FIPS:
        assert product.content_hash(msg) == "sha256:" + digest
    million = b"a" * 1_000_000
    assert product.content_hash(million) == (
        "sha256:cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0")
    rng = np.random.default_rng(7)
    for n in [1, 55, 56, 57, 63, 64, 65, 119, 120, 127, 128, 129, 1000, 65536 + 3]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert product.content_hash(data) == "sha256:" + hashlib.sha256(data).hexdigest(), n


@pytest.mark.parametrize("focal", [None, 550.4])
def test_request_hash_is_sha256_of_the_request_bytes(product, focal):
    frames = _frames() + [scenes.to_rgba(scenes.plane_frame(200, 190, 60.0, 0.0, 4.0))]
    o = product.reconstruction_options(focal_length_pixels=focal)
    assert product.request_content_hash(frames, o) == _expected(frames, o)
    assert product.request_content_hash([], o) == _expected([], o)


def test_request_identity_covers_frames_order_and_options(product):
    """request_identity_covers_frames_order_and_options, tests/reconstruction_stages.rs:459."""
    frames = _frames()
    h = product.request_content_hash(frames)
    assert h == product.request_content_hash([f.copy() for f in frames])
    assert h != product.request_content_hash(_frames(seed=4)), "pixels"
    assert h != product.request_content_hash([frames[1], frames[0]] + frames[2:]), "order"
    one = [f.copy() for f in frames]
    one[3][0, 0, 0] ^= 1
    assert h != product.request_content_hash(one), "one pixel value"
    o = product.reconstruction_options()
    o.max_dense_working_set_bytes //= 2
    assert h != product.request_content_hash(frames, o), "options"


def test_request_hash_polls_per_64kib_and_cancels(product):
    # 200 x 190 RGBA = 152,000 bytes: three chunks of 64 KiB per frame
    frames = [scenes.to_rgba(scenes.plane_frame(200, 190, 60.0, c, 4.0)) for c in (0.0, 0.2)]
    polls = []
    product.request_content_hash(frames, control=lambda: polls.append(1) and False)
    assert len(polls) == 6
    for at in (1, 4):
        polls = []
        with pytest.raises(product.Vt3dError, match=r"\(-6\)"):  # VT3D_E_CANCELED
            product.request_content_hash(frames, control=lambda: polls.append(1) or len(polls) == at)
        assert len(polls) == at
