"""Writes the parameter-echo sections of the reference's published example outputs
(examples/*/reference/*.mcout, from the SYSTEM block to the line before "Time to read
parameters") as fixtures next to the example inputs: tests/golden/examples/<name>_echo.txt.
Run in a checkout that has the reference at /root/reference; the fixtures are data only."""
import os

REF = "/root/reference/examples"
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples")
CASES = {
    "ex01_cubic": "example01/reference/cubic.mcout",
    "ex02_cubic": "example02/reference/cubic.mcout",
    "ex03_fef3": "example03/reference/fef3.mcout",
    "ex04_bmno_pt": "example04/reference/pt_with_uniform_temperatures/bmno.mcout",
    "ex05_snwo": "example05/reference/snwo.mcout",
}


def echo_section(lines):
    i = next(k for k, l in enumerate(lines) if l.strip() == "SYSTEM") - 2
    j = next(k for k, l in enumerate(lines) if l.startswith(" Time to read parameters"))
    return lines[i:j]


if __name__ == "__main__":
    for name, rel in CASES.items():
        with open(os.path.join(REF, rel)) as f:
            sec = echo_section(f.read().splitlines())
        with open(os.path.join(HERE, name + "_echo.txt"), "w") as f:
            f.write("\n".join(sec) + "\n")
        print(name, len(sec), "lines")
