"""Locations of the in-tree native libraries."""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "lib")


def native_lib() -> str:
    return os.path.join(LIB_DIR, "libngpb.so")


def workload_lib() -> str:
    return os.path.join(LIB_DIR, "libngpb_workload.so")


def oracle_lib() -> str:
    return os.path.join(REPO_DIR, "oracle", "_build", "libngpb_oracle.so")
