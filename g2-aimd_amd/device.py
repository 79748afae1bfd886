"""placeholder"""
def hip_lib():
    raise RuntimeError("libg2hip.so not built")
class Executor:
    pass
