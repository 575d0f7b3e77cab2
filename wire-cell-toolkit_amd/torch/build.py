"""Builds the libtorch extension (wct_torch) in-tree, next to the HIP
library it links, with torch.utils.cpp_extension (host compiler only:
the device code is libwct_hip.so)."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
BUILD_DIR = os.path.join(PKG, "_build_torch")


def build(verbose=False):
    from torch.utils.cpp_extension import load
    os.makedirs(BUILD_DIR, exist_ok=True)
    return load(name="wct_torch", sources=[os.path.join(HERE, "wct_torch.cpp")],
                extra_cflags=["-O2", "-std=c++17"],
                extra_ldflags=[f"-L{PKG}", "-lwct_hip", f"-Wl,-rpath,{PKG}"],
                build_directory=BUILD_DIR, verbose=verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(os.path.join(BUILD_DIR, "wct_torch.so"))
