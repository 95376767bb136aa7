"""Host side of the MI355X ReactiveFoam hot path: ctypes bindings of the
drop-in C ABI (libreactiveTransport.so), the HEM model image builder and the
synthetic Cartesian case generator."""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PACKAGE_DIR)
LIBRARY_PATH = os.path.join(ROOT_DIR, "lib", "libreactiveTransport.so")
