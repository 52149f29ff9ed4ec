"""ImageCompare on the GPU: error metrics between two images and an error heat map.

`compare_tensors` runs fhp_image_compare on device images; `compare_files` and the command line
(`python -m falcor_amd.image_compare a.exr b.exr -m mse -t 1e-4 -e heat.png`) follow
Source/Tools/ImageCompare/ImageCompare.cpp:294-420: the images are loaded as RGBA float, must have
the same size, the error is printed, and the comparison passes when the error is finite and at
most the threshold (default 0). Metrics: mse, rmse (relative), mae, mape (ImageCompare.cpp:256-261;
mae squares the differences, as the reference's does).
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import sys

import numpy as np

from . import _lib

METRICS = {"mse": (0, "Mean Squared Error"), "rmse": (1, "Relative Mean Squared Error"),
           "mae": (2, "Mean Absolute Error"), "mape": (3, "Mean Absolute Percentage Error")}


def compare_tensors(a, b, metric: str = "mse", alpha: bool = False, heat_map: bool = False, error_map: bool = None):
    """a, b: [H, W, 4] float32 device tensors. Returns (error, error map [H, W] or None,
    heat map [H, W, 4] or None)."""
    import torch
    if metric not in METRICS:
        raise ValueError(f"unknown error metric '{metric}'")
    if a.shape != b.shape or a.dim() != 3 or a.shape[2] != 4:
        raise ValueError("images must both be [H, W, 4] of the same size")
    if a.dtype != torch.float32 or b.dtype != torch.float32:
        raise ValueError("images must be float32")
    a, b = a.contiguous(), b.contiguous()
    h, w = a.shape[0], a.shape[1]
    want_map = heat_map if error_map is None else (error_map or heat_map)
    emap = torch.empty((h, w), dtype=torch.float32, device=a.device) if want_map else None
    heat = torch.empty((h, w, 4), dtype=torch.float32, device=a.device) if heat_map else None
    err = C.c_double()
    _lib.check(_lib.load().fhp_image_compare(METRICS[metric][0], a.data_ptr(), b.data_ptr(), w, h, int(bool(alpha)),
                                             emap.data_ptr() if emap is not None else None,
                                             heat.data_ptr() if heat is not None else None, C.byref(err),
                                             _lib.stream_ptr()))
    return err.value, emap, heat


def compare_files(path_a, path_b, metric: str = "mse", threshold: float = 0.0, alpha: bool = False,
                  heat_map_path=None, out=sys.stdout, err_out=sys.stderr):
    """compareImages (ImageCompare.cpp:294-345): returns (passed, error or None)."""
    import torch
    from .image_io import read_rgba, save_image
    imgs = []
    for p in (path_a, path_b):
        try:
            imgs.append(read_rgba(p))
        except (OSError, ValueError) as e:
            print(f"Cannot load image from '{p}' (Error: {e}).", file=err_out)
            return False, None
    if imgs[0].shape != imgs[1].shape:
        print("Cannot compare images with different resolutions.", file=err_out)
        return False, None
    dev = torch.device("cuda", torch.cuda.current_device())
    ta, tb = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in imgs)
    error, _, heat = compare_tensors(ta, tb, metric, alpha=alpha, heat_map=heat_map_path is not None)
    if heat is not None:
        save_image(heat_map_path, heat.cpu().numpy())
    print(error, file=out)
    if math.isnan(error) or math.isinf(error):
        return False, error
    return error <= threshold, error


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="ImageCompare", description="Utility to compare images.")
    ap.add_argument("-l", action="store_true", help="List available error metrics.")
    ap.add_argument("-m", default="mse", help="The error metric.")
    ap.add_argument("-t", type=float, default=0.0, help="The error threshold.")
    ap.add_argument("-a", action="store_true", help="Include alpha channel.")
    ap.add_argument("-e", default=None, help="Generate error heat map.")
    ap.add_argument("image1", nargs="?")
    ap.add_argument("image2", nargs="?")
    args = ap.parse_args(argv)
    if args.l:
        print("Available error metrics:")
        for k, (_, desc) in METRICS.items():
            print(f"  {k} - {desc}")
        return 0
    if args.image1 is None or args.image2 is None:
        ap.print_usage(sys.stderr)
        return 1
    if args.m not in METRICS:
        print(f"Unknown error metric '{args.m}'.", file=sys.stderr)
        return 1
    ok, _ = compare_files(args.image1, args.image2, args.m, args.t, args.a, args.e)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
