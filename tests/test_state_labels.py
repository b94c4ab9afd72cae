"""Per-label state accessors named after the reference's generated binding
(harmonize/python/templates.py:85-152, runtime.py:1348-1371): one field of a
program's device state at a time, the empty group/thread labels, and the
reference's error for a pointer store into a non-array field."""

import ctypes

import numpy as np
import pytest

import harmonize_amd as hz

_VP = ctypes.c_void_p


class PhotonStateC(ctypes.Structure):
    """hz_photon_state of include/harmonize_amd.h."""
    _fields_ = [(n, ctypes.c_uint32) for n in ("photon_count", "seed_lo", "seed_hi", "bins", "replicas", "max_split",
                                                "max_events", "regions")] + \
               [(n, ctypes.c_float) for n in ("thickness", "mean_free_path", "albedo", "bin_scale", "region_scale",
                                               "split_above", "roulette_below", "pad")] + \
               [("tally", _VP), ("events", _VP), ("iterator", _VP)]


def fn(name, restype=None, argtypes=(_VP, _VP)):
    f = getattr(ctypes.CDLL(str(hz.LIB_PATH)), name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


def last_error():
    f = fn("hz_last_error", ctypes.c_char_p, ())
    msg = f()
    return msg.decode() if msg else None


def test_pointer_store_into_a_scalar_field_is_the_reference_error():
    # Raised before any device access, so no GPU is needed.
    fn("store_pointer_state_device_bins_photon_async")(None, None)
    assert last_error() == "Attempted to set indirection of non-array state field."
    fn("store_pointer_state_device_collatz_event")(None, None)
    assert last_error() == "Attempted to set indirection of non-array state field."
    fn("store_pointer_state_group_bfs_async")(None, None)
    assert "non-array" in last_error()


def test_empty_group_and_thread_labels():
    fn("load_state_group_collatz_async")(None, None)
    assert last_error() is None
    fn("store_state_thread_barrier_async")(None, None)
    assert last_error() is None


@pytest.mark.gpu
def test_field_accessors_on_the_device_state():
    alloc = fn("alloc_state_collatz_async", _VP, ())
    free = fn("free_state_collatz_async", None, (_VP,))
    dev = alloc()
    assert dev and last_error() is None
    try:
        steps = hz.DeviceArray(16, np.uint32)
        starts = hz.DeviceArray(16, np.uint64)
        # A pointer field set by value, another through the pointer store.
        val = ctypes.c_uint64(starts.ptr)
        fn("store_state_device_starts_collatz_async")(dev, ctypes.addressof(val))
        assert last_error() is None
        fn("store_pointer_state_device_steps_collatz_async")(dev, steps.ptr)
        assert last_error() is None
        fn("store_pointer_state_device_iterator_collatz_async")(dev, 0x1234)
        # Whole-state load agrees with the per-field writes.
        whole = (ctypes.c_uint64 * 3)()
        fn("load_state_device_collatz_async")(ctypes.addressof(whole), dev)
        assert list(whole) == [starts.ptr, steps.ptr, 0x1234]
        out = ctypes.c_uint64(0)
        fn("load_state_device_steps_collatz_async")(ctypes.addressof(out), dev)
        assert out.value == steps.ptr
    finally:
        free(dev)

    # A scalar field of the photon state, at its offset.
    alloc_p = fn("alloc_state_photon_event", _VP, ())
    free_p = fn("free_state_photon_event", None, (_VP,))
    dev = alloc_p()
    try:
        zero = (ctypes.c_uint8 * ctypes.sizeof(PhotonStateC))()
        fn("store_state_device_photon_event")(dev, ctypes.addressof(zero))
        bins = ctypes.c_uint32(1000)
        albedo = ctypes.c_float(0.9)
        fn("store_state_device_bins_photon_event")(dev, ctypes.addressof(bins))
        fn("store_state_device_albedo_photon_event")(dev, ctypes.addressof(albedo))
        st = PhotonStateC()
        fn("load_state_device_photon_event")(ctypes.addressof(st), dev)
        assert st.bins == 1000 and abs(st.albedo - 0.9) < 1e-7 and st.photon_count == 0
    finally:
        free_p(dev)


@pytest.mark.gpu
def test_wall_clock_rate():
    rate = fn("wall_clock_rate", ctypes.c_longlong, ())()
    assert rate >= 1_000_000  # MI355X's constant clock is 100 MHz
