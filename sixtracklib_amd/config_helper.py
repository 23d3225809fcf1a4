"""python/sixtracklib/config_helper.py ``supports``: the architectures this
build tracks on.  "hip" and "cuda" (the reference's GPU job, which this
package runs on HIP) are supported; "cpu" and "opencl" are not."""


def supports(architecture):
    return str(architecture).strip().lower() in ("hip", "cuda")
