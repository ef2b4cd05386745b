"""The render manager's command queue system without a GPU
(include/magik_manager.h): its declarations against the reference's magik.h
(sizes, offsets and enum values, through oracle/_ref), and the push /
dispatch / fetch rules of command.cpp:683-737 and framebuffer.cpp:6-52.
Without a device the worker thread stops at once and records the error;
the command queue and the collections still behave as on a GPU."""
import ctypes as C
import json
import os
import time

import pytest

import magik_amd as M
from magik_amd import abi
from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))


def test_manager_declarations_match_magik_h():
    buf = (C.c_size_t * 256)()
    n = O.lib().orc_manager_layout(buf)
    assert list(buf[:n]) == GOLD["manager_layout"]
    assert C.sizeof(abi.ManagerDescriptor) == GOLD["manager_layout"][0]
    assert C.sizeof(abi.RendererStatistics) == GOLD["manager_layout"][7]
    assert C.sizeof(abi.HostContainer) == GOLD["manager_layout"][14]


@pytest.fixture
def cpu_manager():
    import torch
    if torch.cuda.is_available():
        pytest.skip("runs where no GPU is visible")
    m = M.Manager(0, n_chunk=6)
    yield m
    m.close()


def test_worker_without_a_device_records_the_error(cpu_manager):
    for _ in range(200):
        if cpu_manager.error() != abi.MAGIK_SUCCESS:
            break
        time.sleep(0.01)
    assert cpu_manager.error() == 4            # MAGIK_INVALID_CUDA_DEVICE
    r, new = cpu_manager.fetch()
    assert r == abi.MAGIK_SUCCESS and not new  # not running: nothing to fetch
    assert cpu_manager.extract("test") is None


def test_push_rules(cpu_manager):
    m = cpu_manager
    assert m.push(M.Manager.command("kerr_spin", 0.5)) == abi.MAGIK_SUCCESS                   # 2 chunks
    assert m.push(M.Manager.command("camera_resolution", 64, 48)) == abi.MAGIK_SUCCESS      # 3 chunks
    assert m.push(M.Manager.command("abs_tol", 1e-4)) == 402                                 # 7 > 6 chunks: overflow
    assert m.push(M.Manager.command("render")) == abi.MAGIK_SUCCESS                          # 6 of 6
    bad = (C.c_uint32 * 2)(12345, 0)
    assert m.push(bad) == 403                                                                # not a command
    raw = (C.c_uint8 * 16)()
    C.memmove(raw, M.Manager.command("render"), 4)
    assert m.lib.magik_cqs_push_command(m.h, C.c_void_p(C.addressof(raw) + 1)) == 404          # misaligned
    assert m.lib.magik_cqs_push_command(m.h, None) == abi.MAGIK_ERROR_INVALID_POINTER


def test_dropped_overflow_and_dispatch_handshake():
    import torch
    if torch.cuda.is_available():
        pytest.skip("runs where no GPU is visible")
    m = M.Manager(0, n_chunk=2, drop_overflow=True)
    try:
        assert m.push(M.Manager.command("kerr_spin", 0.5)) == abi.MAGIK_SUCCESS
        assert m.push(M.Manager.command("kerr_spin", 0.6)) == abi.MAGIK_SUCCESS   # dropped, not an error
        assert m.dispatch() is True        # the worker's buffer starts consumed
        assert m.dispatch() is False       # the worker has not consumed it (no device): no swap
    finally:
        m.close()


def test_collections():
    lib = abi.load()
    fb = C.c_void_p()
    for cfg in (abi.FB_CONFIG_OPENGL, abi.FB_CONFIG_VULKAN, abi.FB_CONFIG_CUDA, 7):
        assert lib.magik_configure_framebuffer_collection(C.byref(fb), cfg) == abi.MAGIK_UNKNOWN_ENUM_TYPE
    assert lib.magik_configure_framebuffer_collection(C.byref(fb), abi.FB_CONFIG_HOST) == abi.MAGIK_SUCCESS
    assert lib.magik_destroy_framebuffer_collection(fb) == abi.MAGIK_SUCCESS
    assert lib.magik_configure_framebuffer_collection(None, abi.FB_CONFIG_HOST) == abi.MAGIK_ERROR_INVALID_POINTER
    st = abi.RendererStatistics(x_sample=5)
    assert lib.magik_destroy_renderer_statistics(C.byref(st)) == abi.MAGIK_SUCCESS and st.x_sample == 0
    # a fetch hands out the six histories as one host block that destroy frees
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    block = libc.malloc(6 * 3 * 4)
    st = abi.RendererStatistics(latest_wavefront_history_length=3)
    for i, q in enumerate(abi.QUEUE_HISTORIES):
        setattr(st, f"h_latest_{q}_queue_count_history", block + 12 * i)
    assert lib.magik_destroy_renderer_statistics(C.byref(st)) == abi.MAGIK_SUCCESS
    assert st.latest_wavefront_history_length == 0 and st.h_latest_splat_queue_count_history is None
    d = abi.ManagerDescriptor(display_type=9)
    h = C.c_void_p()
    assert lib.magik_render_manager_create(C.byref(h), d) == abi.MAGIK_UNKNOWN_ENUM_TYPE and not h.value
