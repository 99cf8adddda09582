"""Drop-in for the reference's Python module `libakaze_pybindings`
(python/python_bindings.cpp:27-40): the same three classes and methods, so
that `import akaze_amd.libakaze_pybindings as akaze` replaces
`import libakaze_pybindings as akaze` in a caller such as python/test.py.

  AKAZEOptions()            setWidth(w), setHeight(h)
  AKAZE(options)            Create_Nonlinear_Scale_Space(float32 H x W in [0,1]) -> 0 / -1
                            Feature_Detection()   -> n x 7 float32
                            Compute_Descriptors() -> (desc n x 61 uint8, kpts n x 7 float32)
  Matcher()                 BFMatch(desc_query, desc_train) -> n x 8 float32

Keypoint rows are [x, y, size, angle, response, octave, class_id]
(mat2kpvec, src/lib/AKAZE.cpp:290-298); BFMatch rows are
[q, t1, 0, d1, q, t2, 0, d2] (src/lib/AKAZE.cpp:82-105).  Every call runs the
HIP library; errors raise akaze_amd.AkazeError.
"""
import akaze_amd as _A

__all__ = ["AKAZEOptions", "AKAZE", "Matcher"]


class AKAZEOptions:
    """AKAZEOptions with the reference defaults (AKAZEConfig.h:86-111)."""

    def __init__(self):
        self._o = _A.AKAZEOptions()

    def setWidth(self, w):
        self._o.setWidth(int(w))

    def setHeight(self, h):
        self._o.setHeight(int(h))


class AKAZE:
    """One image size per instance, fixed by the options at construction."""

    def __init__(self, options):
        # the reference copies the options (class_<AKAZE>(init<AKAZEOptions>))
        o = _A.AKAZEOptions()
        for name, _ in _A.COptions._fields_:
            setattr(o, name, getattr(options._o, name))
        self._ak = _A.AKAZE(o)

    def Create_Nonlinear_Scale_Space(self, img):
        return self._ak.Create_Nonlinear_Scale_Space(img)

    def Feature_Detection(self):
        return self._ak.Feature_Detection()

    def Compute_Descriptors(self):
        return self._ak.Compute_Descriptors()


class Matcher:
    def __init__(self):
        self._m = _A.Matcher()

    def BFMatch(self, desc_query, desc_train):
        return self._m.bfmatch_(desc_query, desc_train)
