"""The M-LDB variants on the GPU against the oracle, bit for bit: upright
(descriptor 4: angle 0, no orientation), the subsampled descriptor of
descriptor_size > 0, one and two channels, single-image and batch contexts
(graph replay), and the matcher on the shorter rows."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

VARIANTS = [dict(descriptor=4), dict(descriptor_size=256), dict(descriptor_size=64),
            dict(descriptor_channels=1), dict(descriptor_channels=2),
            dict(descriptor=4, descriptor_size=128, descriptor_channels=2), dict(descriptor_size=100,
                                                                                 descriptor_channels=1),
            dict(descriptor_size=486)]


def name(kw):
    return ",".join("%s=%s" % kv for kv in sorted(kw.items()))


@pytest.mark.parametrize("kw", VARIANTS, ids=[name(k) for k in VARIANTS])
def test_variant_single_image_matches_oracle(gpu_lib, kw):
    A = gpu_lib
    W, H = 640, 480
    img = A.to_float(A.synth_image(31, W, H))
    ak = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H, **kw))
    kps, desc = ak.run(img)
    okp, odesc = O.Pipeline(W, H, **kw).run(img)
    nbits = kw.get("descriptor_size", 0) or 162 * kw.get("descriptor_channels", 3)
    print("%s: %d keypoints, %d bytes per descriptor" % (name(kw), len(kps), desc.shape[1]), flush=True)
    assert desc.shape[1] == (nbits + 7) // 8 == odesc.shape[1]
    assert len(kps) == len(okp) > 100
    assert kps.tobytes() == okp.tobytes()
    assert np.array_equal(desc, odesc)
    if kw.get("descriptor") == 4:
        assert not kps["angle"].any()
    # matcher on the shorter rows, against the oracle's brute force
    m = A.Matcher().match2nn(desc, desc[::-1].copy())
    om = O.match2nn(desc, desc[::-1].copy())
    assert np.array_equal(m["trainIdx"], om["trainIdx"]) and np.array_equal(m["distance"], om["distance"])


def test_variants_in_a_batch_graph(gpu_lib):
    A = gpu_lib
    W, H = 1920, 1080
    kw = dict(descriptor_size=256, descriptor_channels=2)
    imgs = [A.to_float(A.synth_image(41 + z, W, H)) for z in range(3)]
    ref = [O.Pipeline(W, H, **kw).run(im) for im in imgs]
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H, **kw), 3)
    bk.upload(imgs)
    for graph in (False, True, True):
        bk.set_graph_mode(graph)
        bk.process_async(3)
        cnt, _ = bk.counts(3)
        for z in range(3):
            kps, desc, n = bk.read_results(z, int(cnt[z]))
            assert kps.tobytes() == ref[z][0].tobytes(), (graph, z)
            assert np.array_equal(desc, ref[z][1]), (graph, z)


def test_unsupported_descriptors_are_refused(gpu_lib):
    A = gpu_lib
    for kw, status in ((dict(descriptor=1), 6), (dict(descriptor_channels=4), 1), (dict(descriptor_size=487), 1),
                       (dict(descriptor_size=163, descriptor_channels=1), 1)):
        with pytest.raises(A.AkazeError) as e:
            A.AKAZE(A.AKAZEOptions(img_width=640, img_height=480, **kw))
        assert e.value.status == status, kw
