"""GPU parity of same-size image batches (grid z = image) against the CPU
oracle, image by image: keypoints and descriptors must be identical to the
oracle's for each image, whatever the other images of the batch hold."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def oracle_run(W, H, img, **kw):
    orc = O.Pipeline(W, H, **kw)
    return orc.run(img)


def check_image(bk, z, okp, odesc, where):
    kps, desc, n = bk.read_results(z, len(okp) + 16)
    assert n == len(okp), (where, z, n, len(okp))
    assert kps.tobytes() == okp.tobytes(), (where, z)
    assert np.array_equal(desc, odesc), (where, z)


def test_batch_eager_mixed_density(gpu_lib):
    """Four 640x480 images in one batch, one of them noise with tens of
    thousands of candidates (global-memory suppression path) beside sparse
    ones (LDS path); then a smaller batch reuses the context."""
    A = gpu_lib
    W, H, thr = 640, 480, 2e-4
    rng = np.random.default_rng(3)
    imgs = [A.to_float(A.synth_image(10 + i, W, H)) for i in range(3)]
    imgs.insert(2, rng.random((H, W)).astype(np.float32))
    ref = [oracle_run(W, H, im, dthreshold=thr) for im in imgs]
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H, dthreshold=thr), 4)
    bk.upload(imgs)
    bk.process_async(4)
    counts, st = bk.counts(4)
    assert list(counts) == [len(r[0]) for r in ref]
    assert all(s == 0 for s in st)
    for z in range(4):
        check_image(bk, z, ref[z][0], ref[z][1], "batch4")
    # level readback of image 2 (the noise image)
    orc = O.Pipeline(W, H, dthreshold=thr)
    orc.scale_space(imgs[2])
    for lvl in (0, 5, bk.nlev - 1):
        assert np.array_equal(bk.read_level(2, lvl, 0), orc.level(lvl, 0)), lvl
    imgs2 = [A.to_float(A.synth_image(40 + i, W, H)) for i in range(2)]
    ref2 = [oracle_run(W, H, im, dthreshold=thr) for im in imgs2]
    bk.upload(imgs2)
    bk.process_async(2)
    for z in range(2):
        check_image(bk, z, ref2[z][0], ref2[z][1], "batch2")


def test_batch_graph_sizes(gpu_lib):
    """Graph mode captures one graph per batch size; replays of sizes 3, 1,
    3 with eager work in between each give the oracle's results."""
    A = gpu_lib
    W, H = 320, 240
    imgs = [A.to_float(A.synth_image(60 + i, W, H)) for i in range(3)]
    ref = [oracle_run(W, H, im) for im in imgs]
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 3)
    bk.set_graph_mode(True)
    bk.upload(imgs)
    other = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H))
    for rep, n in enumerate([3, 1, 3, 1]):
        bk.process_async(n)
        bk.wait()
        for z in range(n):
            check_image(bk, z, ref[z][0], ref[z][1], "graph rep %d" % rep)
        other.run(A.to_float(A.synth_image(200 + rep, W, H)))


def test_batch_1080p_equals_single(gpu_lib):
    """Two 1920x1080 images: the batch result of each equals a single-image
    context's result and the oracle's."""
    A = gpu_lib
    W, H = 1920, 1080
    imgs = [A.to_float(A.synth_image(70 + i, W, H)) for i in range(2)]
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 2)
    bk.upload(imgs)
    bk.process_async(2)
    bk.wait()
    for z in range(2):
        single = A.AKAZE(A.AKAZEOptions(img_width=W, img_height=H))
        skp, sdesc = single.run(imgs[z])
        okp, odesc = oracle_run(W, H, imgs[z])
        assert skp.tobytes() == okp.tobytes() and np.array_equal(sdesc, odesc)
        check_image(bk, z, okp, odesc, "1080p")


def test_ragged_batches_match_oracle(gpu_lib):
    """Images of three sizes in one ragged call (one batch per size class on
    concurrent streams), eager and graph mode: each image's result is the
    oracle's."""
    A = gpu_lib
    sizes = [(320, 240), (640, 480), (500, 380)]
    order = [0, 1, 2, 1, 0, 2, 1]
    imgs = [A.to_float(A.synth_image(80 + i, *sizes[k])) for i, k in enumerate(order)]
    ref = [oracle_run(im.shape[1], im.shape[0], im) for im in imgs]
    rg = A.Ragged(A.AKAZEOptions(), sizes, [3, 3, 2])
    for graph in (False, True):
        rg.set_graph_mode(graph)
        for rep in range(2 if graph else 1):
            res = rg.process(imgs)
            for i, (kps, desc) in enumerate(res):
                assert kps.tobytes() == ref[i][0].tobytes(), (graph, rep, i)
                assert np.array_equal(desc, ref[i][1]), (graph, rep, i)


def test_ragged_five_size_classes_match_oracle(gpu_lib):
    """The five size classes of the ragged configuration (320x240 ...
    2560x1440) in one ragged call, eager and then replayed from graphs: each
    image's keypoints and descriptors are the oracle's."""
    A = gpu_lib
    sizes = [(320, 240), (640, 480), (1280, 720), (1920, 1080), (2560, 1440)]
    order = [4, 2, 3, 0, 1, 3, 2, 4, 1]
    imgs = [A.to_float(A.synth_image(120 + i, *sizes[k])) for i, k in enumerate(order)]
    ref = [oracle_run(im.shape[1], im.shape[0], im) for im in imgs]
    caps = [order.count(k) for k in range(len(sizes))]
    rg = A.Ragged(A.AKAZEOptions(), sizes, caps)
    # per-class batch calls, then the persistent work-list kernels over the
    # whole call (eager, then one captured graph replayed), then a different
    # class composition (a subset of the images) from a second graph
    for lists in (False, True):
        rg.set_work_lists(lists)
        for graph in (False, True):
            rg.set_graph_mode(graph)
            for rep in range(2 if graph else 1):
                print("ragged 5 classes: work lists %s graph %s rep %d" % (lists, graph, rep), flush=True)
                res = rg.process(imgs)
                for i, (kps, desc) in enumerate(res):
                    assert len(kps) > 0
                    assert kps.tobytes() == ref[i][0].tobytes(), (lists, graph, rep, i, sizes[order[i]])
                    assert np.array_equal(desc, ref[i][1]), (lists, graph, rep, i, sizes[order[i]])
        sub = [0, 2, 3, 6]
        res = rg.process([imgs[i] for i in sub])
        for j, i in enumerate(sub):
            assert res[j][0].tobytes() == ref[i][0].tobytes() and np.array_equal(res[j][1], ref[i][1]), (lists, i)


def test_every_fed_variant_and_the_autotuner(gpu_lib):
    """Each blocked FED kernel variant (ids 2..19: LDS regions, register
    regions, wavefront strips) on every level, and the autotuner's choice,
    give the oracle's levels bit for bit."""
    A = gpu_lib
    W, H = 900, 675
    img = A.to_float(A.synth_image(91, W, H))
    orc = O.Pipeline(W, H)
    okp, odesc = orc.run(img)
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 2)
    for v in range(2, 20):
        bk.set_fed_variants([v] * bk.nlev)
        bk.upload([img, img])
        bk.process_async(2)
        for z in range(2):
            check_image(bk, z, okp, odesc, "variant %d" % v)
        assert np.array_equal(bk.read_level(1, bk.nlev - 1, 0), orc.level(bk.nlev - 1, 0)), v
    chosen = bk.autotune(2, reps=2)
    assert all(2 <= v <= 19 for v in chosen[1:])
    bk.upload([img, img])
    bk.process_async(2)
    check_image(bk, 1, okp, odesc, "autotuned %s" % list(chosen))


def test_strip_variants_on_exactly_halving_levels(gpu_lib):
    """The wavefront-strip FED variants (17: LDS first launch, 18: strips
    with smooth + flow, 19: the same with 16 steps per launch) on sizes
    whose octaves halve exactly, where the first launch of an octave forms
    the halved input while loading, with one and with three images per
    launch; every level against the oracle's."""
    A = gpu_lib
    for (W, H, seed) in [(640, 480, 97), (1280, 720, 98)]:
        img = A.to_float(A.synth_image(seed, W, H))
        orc = O.Pipeline(W, H)
        okp, odesc = orc.run(img)
        for nz in (1, 3):
            bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), nz)
            for v in (17, 18, 19):
                bk.set_fed_variants([v] * bk.nlev)
                bk.upload([img] * nz)
                bk.process_async(nz)
                for z in range(nz):
                    check_image(bk, z, okp, odesc, "variant %d %dx%d nz %d" % (v, W, H, nz))
                for lvl in range(1, bk.nlev):
                    assert np.array_equal(bk.read_level(nz - 1, lvl, 0), orc.level(lvl, 0)), (v, W, H, nz, lvl)
                    assert np.array_equal(bk.read_level(nz - 1, lvl, 1), orc.level(lvl, 1)), (v, W, H, nz, lvl)


def test_determinant_images_on_request(gpu_lib):
    """The asynchronous path writes only the candidates' determinant
    neighbourhoods unless asked: the keypoints are the same either way, the
    full images come back when kept, and a readback that would see partial
    images is refused."""
    A = gpu_lib
    W, H = 640, 480
    img = A.to_float(A.synth_image(95, W, H))
    orc = O.Pipeline(W, H)
    okp, odesc = orc.run(img)
    bk = A.BatchAKAZE(A.AKAZEOptions(img_width=W, img_height=H), 1)
    bk.upload([img])
    for keep in (False, True, False):
        bk.set_keep_determinant(keep)
        bk.process_async(1)
        check_image(bk, 0, okp, odesc, "keep %s" % keep)
        if keep:
            for lvl in range(bk.nlev):
                assert np.array_equal(bk.read_level(0, lvl, 4), orc.level(lvl, 4)), lvl
        else:
            with pytest.raises(A.AkazeError) as e:
                bk.read_level(0, 0, 4)
            assert e.value.status == 5
