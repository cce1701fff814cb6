"""Display post (the `Post` class of src/demo/post.js:131-210) over the C ABI of include/post.h.

    post = Post(0, "bgra8unorm", bloom=0.03, vignette=0.35)
    target = post.encode(left=hush_out, right=svgf_out, split=0.5, exposure=1.0)   # 8-bit Surface

Images are rgba16float `Surface`s or (H, W, 4) float16 torch tensors on the GPU; the target is a Surface of
4-byte texels (allocated here when not given) or an (H, W, 4) uint8 tensor.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hip import Surface

FORMATS = {"rgba8unorm": 0, "bgra8unorm": 1}
SURFACES = {"hdr": 0, **{f"level{i}": i for i in range(1, 7)}}


class PostError(Exception):
    def __init__(self, message: str, code: int = -1):
        super().__init__(message)
        self.code = code


class Post:
    """new Post(device, format, {bloom = 0.03, vignette = 0.35}) (post.js:131-147)."""

    def __init__(self, device=0, format: str = "bgra8unorm", opts=None, **kw):
        self._L = _lib.load()
        if format not in FORMATS:
            raise PostError(f"Post: unsupported target format '{format}' (rgba8unorm, bgra8unorm)")
        self.format = format
        o = _lib.post_options()
        self._L.post_default_options(ctypes.byref(o))
        given = dict(opts or {})
        given.update(kw)
        for k, v in given.items():
            if k not in ("bloom", "vignette"):
                raise PostError(f"Post: unknown option '{k}'")
            setattr(o, k, float(v))
        self.bloom, self.vignette = float(o.bloom), float(o.vignette)
        p = ctypes.c_void_p()
        r = self._L.post_create(int(getattr(device, "ordinal", device)), ctypes.byref(o), ctypes.byref(p))
        if r != 0:
            raise PostError(self._L.post_last_error(None).decode(), r)
        self._h = p.value
        self._target = None

    def _check(self, r: int) -> None:
        if r != 0:
            raise PostError(self._L.post_last_error(ctypes.c_void_p(self._h)).decode(), r)

    @staticmethod
    def _image(t) -> _lib.hush_surface:
        from . import HushError, _surface_struct
        try:
            return _surface_struct(t)
        except HushError as e:
            raise PostError(str(e).replace("Hush.denoise", "Post.encode")) from None

    def _target_of(self, target, width, height):
        if target is None:
            if self._target is None or (self._target.width, self._target.height) != (width, height):
                self._target = Surface(width, height, np.uint8, 4)
            target = self._target
        if isinstance(target, Surface):
            return target, target.ptr, target.pitch, target.width, target.height
        if hasattr(target, "data_ptr"):   # torch uint8 (H, W, 4)
            if target.dim() != 3 or target.shape[2] != 4 or target.stride(2) != 1 or target.stride(1) != 4 \
                    or str(target.dtype) != "torch.uint8":
                raise PostError("Post.encode: a tensor target must be (H, W, 4) uint8 with contiguous texels")
            return target, target.data_ptr(), target.stride(0), target.shape[1], target.shape[0]
        raise PostError(f"Post.encode: unsupported target type {type(target).__name__}")

    def encode(self, left, right, mid=None, split: float = 0.5, split2: float | None = None, same: bool = False,
               exposure: float = 1.0, target=None, width: int | None = None, height: int | None = None,
               line: float = 1.5, stream: int | None = None):
        """encode(encoder, {left, right, mid = right, split = 0.5, split2 = split, same = false, exposure = 1,
        target, width, height, line = 1.5}) (post.js:179-209); returns the target."""
        a = _lib.post_args()
        a.left, a.right = self._image(left), self._image(right)
        if mid is not None:
            a.mid = self._image(mid)
        w = width if width is not None else a.left.width
        h = height if height is not None else a.left.height
        tgt, ptr, pitch, tw, th = self._target_of(target, w, h)
        a.split = float(split)
        a.split2 = float(split if split2 is None else split2)
        a.same = int(bool(same))
        a.exposure = float(exposure)
        a.line = float(line)
        a.target, a.target_pitch, a.width, a.height = ptr, pitch, tw, th
        a.format = FORMATS[self.format]
        self._check(self._L.post_encode(self._h, ctypes.byref(a), ctypes.c_void_p(stream or 0)))
        return tgt

    def resize(self, width: int, height: int) -> None:
        self._check(self._L.post_resize(self._h, int(width), int(height)))

    def surface(self, name) -> Surface:
        """Parity hook: the composited HDR image ("hdr") or a bloom level ("level1" .. "level6")."""
        s = _lib.hush_surface()
        self._check(self._L.post_debug_surface(self._h, SURFACES[name] if isinstance(name, str) else int(name),
                                               ctypes.byref(s)))
        return Surface(s.width, s.height, np.float16, 4, ptr=s.ptr, pitch=s.pitch, owner=self)

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            self._L.post_destroy(self._h)
            self._h = None
        self._target = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass
