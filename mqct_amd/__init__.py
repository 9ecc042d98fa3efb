"""MI355X-native implementation of MQCT's two data-parallel hot paths: the
state-to-state matrix build and the quantum amplitude propagation, behind
the C ABI of libmqct_hip.so (include/mqct_hip.h)."""
__version__ = "0.1.0"
