"""GPU runtime: the ctypes binding of libxentica_hip.so, device arrays and step engines."""
