"""Python side of the MI355X-native CAVIAR2 MD step: ctypes bindings to the
C-ABI library (libcv2md.so) and synthetic system builders."""
