"""Generate sax_bp256.inc: the card-256 iSAX breakpoints float32(Phi^-1(i/256)),
i = 1..255, from the closed form (the reference's table slice at offset
255*254/2 holds exactly these values).  Run: python gen_breakpoints.py"""
import os
import struct
from statistics import NormalDist

def main() -> None:
    vals = []
    for i in range(1, 256):
        x = NormalDist().inv_cdf(i / 256) if i != 128 else 0.0
        f = struct.unpack("<f", struct.pack("<f", x))[0]  # round to float32
        vals.append(f)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "sax_bp256.inc"), "w") as fh:
        fh.write("/* generated by gen_breakpoints.py: float32(Phi^-1(i/256)), i=1..255 */\n")
        for i in range(0, 255, 4):
            fh.write("  " + ", ".join(f"{v!r}f" if "e" in repr(v) or "." in repr(v) else f"{v!r}.0f"
                                     for v in vals[i:i + 4]) + ",\n")

if __name__ == "__main__":
    main()
