"""zaptex_encode_dev_noalloc: the encoders on the library's per-device scratch array give the bytes of
zaptex_encode_dev with a caller workspace, call after call (the second call smaller than the first, so stale
lists and counters from the first are in the array), and images that need more than the array are refused."""
import numpy as np
import pytest

from zap_amd import binding, synth

pytestmark = pytest.mark.gpu


def _images(zt, specs):
    """specs: (image HxWx4 numpy, fmt) -> (device sources, reference outputs, no-alloc outputs, image tables)"""
    import torch
    srcs, outs_a, outs_b = [], [], []
    for img, fmt in specs:
        h, w = img.shape[:2]
        srcs.append(torch.from_numpy(np.ascontiguousarray(img)).cuda())
        n = binding.bc_size(w, h, fmt)
        outs_a.append(torch.full((n,), 0xA5, dtype=torch.uint8, device="cuda"))
        outs_b.append(torch.full((n,), 0x5A, dtype=torch.uint8, device="cuda"))
    tab = lambda outs: zt.images((s, o, img.shape[1], img.shape[0], img.shape[1] * img.itemsize * 4, fmt)
                                 for s, o, (img, fmt) in zip(srcs, outs, specs))
    return srcs, outs_a, outs_b, tab(outs_a), tab(outs_b)


def _check_call(zt, specs):
    import torch
    srcs, outs_a, outs_b, imgs_a, imgs_b = _images(zt, specs)
    need = zt.workspace_size(imgs_a)
    assert 0 < need <= zt.NOALLOC_SCRATCH_BYTES
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    zt.encode_dev(imgs_a, ws)
    zt.encode_dev_noalloc(imgs_b)
    torch.cuda.synchronize()
    for (img, fmt), a, b in zip(specs, outs_a, outs_b):
        assert torch.equal(a, b), (img.shape, fmt)
    return [b.cpu().numpy() for b in outs_b]


def test_noalloc_matches_workspace_entry_point_twice(zt, oracle):
    c2 = synth.c2_bc7(1024, 2)
    first = [(np.ascontiguousarray(c2[:384, :512]), "bc7"), (synth.c1_bc1(300, 1)[:200], "bc1"),
             (np.ascontiguousarray(c2[:67, :129]), "bc3"), (synth.normal_map(64, 64, 11), "bc5"),
             (np.ascontiguousarray(c2[:3, :5]), "bc4"), (np.ascontiguousarray(c2[:256, :256]), "astc"),
             (synth.c3_hdr(64, 5), "bc6h")]
    got = _check_call(zt, first)
    # the oracle agrees on the smallest images of the first call
    assert np.array_equal(got[2], oracle.bc_encode(first[2][0], "bc3"))
    assert np.array_equal(got[4], oracle.bc_encode(first[4][0], "bc4"))
    # a smaller second call: the array still holds the first call's lists, counters and climb state
    second = [(np.ascontiguousarray(c2[512:608, 512:640]), "bc7"), (np.ascontiguousarray(c2[:64, :64]), "bc1"),
              (np.ascontiguousarray(c2[100:164, 100:164]), "astc")]
    got = _check_call(zt, second)
    assert np.array_equal(got[0], oracle.bc_encode(second[0][0], "bc7"))


def test_noalloc_rejects_images_beyond_the_scratch(zt):
    import torch
    w = h = 2048   # 262,144 BC7 blocks need about 111 MB of workspace
    src = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    dst = torch.zeros(binding.bc_size(w, h, "bc7"), dtype=torch.uint8, device="cuda")
    imgs = zt.images([(src, dst, w, h, w * 4, "bc7")])
    assert zt.workspace_size(imgs) > zt.NOALLOC_SCRATCH_BYTES
    with pytest.raises(binding.ZapTexError, match="scratch"):
        zt.encode_dev_noalloc(imgs)
