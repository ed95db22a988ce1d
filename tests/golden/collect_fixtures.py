"""Collect the reference's own golden data into small committed fixtures.

Everything written here is DATA the reference's tests hold (JSON fixture files,
or numbers quoted inline in its unit tests), never source text.  Run in a
container that has the reference checkout; the fixtures are committed so tests
never read the reference at run time:

    python tests/golden/collect_fixtures.py /root/reference
"""
import json
import shutil
import sys
from pathlib import Path

OUT = Path(__file__).resolve().parent


def main(ref: str):
    r = Path(ref)
    # Boys: scipy hyp1f1 values (tests/golden/boys/reference.json)
    shutil.copy(r / "tests/golden/boys/reference.json", OUT / "boys_reference.json")
    # Rys: RDK roots/weights cross-validation cases (tests/golden/rys/reference.json)
    shutil.copy(r / "tests/golden/rys/reference.json", OUT / "rys_reference.json")
    # PySCF 2.11.0 STO-3G preset systems with S, H, compressed ERI (content/presets/systems/*.json)
    pre = OUT / "presets"
    pre.mkdir(exist_ok=True)
    for f in sorted((r / "content/presets/systems").glob("*.json")):
        shutil.copy(f, pre / f.name)
    # IQCP's own energies (tight, DIIS, 75/Standard/SG-1 grid): tests/benchmarks/criterion/iqcp_iteration_counts.json
    shutil.copy(r / "tests/benchmarks/criterion/iqcp_iteration_counts.json", OUT / "iqcp_energies.json")
    # PySCF energies (tests/golden/dft/validation_all_bases.json)
    shutil.copy(r / "tests/golden/dft/validation_all_bases.json", OUT / "pyscf_energies.json")
    # PySCF gradients (tests/golden/dft/gradient_validation.json)
    shutil.copy(r / "tests/golden/dft/gradient_validation.json", OUT / "pyscf_gradients.json")

    # Geometries (bohr) of the 6 benchmark molecules: crates/qc-core/examples/collect_iteration_counts.rs:36-97
    geoms = {
        "H2": [[1, [0.0, 0.0, 0.0]], [1, [0.0, 0.0, 1.4]]],
        "HF": [[1, [0.0, 0.0, 0.0]], [9, [0.0, 0.0, 1.7328]]],
        "H2O": [[8, [0.0, 0.0, 0.2217282]], [1, [0.0, 1.4305447, -0.8869128]], [1, [0.0, -1.4305447, -0.8869128]]],
        "NH3": [[7, [0.0, 0.0, 0.219705]], [1, [0.0, 1.7714918, -0.512645]],
                [1, [1.5342036, -0.8857459, -0.512645]], [1, [-1.5342036, -0.8857459, -0.512645]]],
        "CH4": [[6, [0.0, 0.0, 0.0]], [1, [1.1851, 1.1851, 1.1851]], [1, [-1.1851, -1.1851, 1.1851]],
                [1, [-1.1851, 1.1851, -1.1851]], [1, [1.1851, -1.1851, -1.1851]]],
        "C6H6": [[6, [0.0, 2.6399473960, 0.0]], [6, [2.2861906655, 1.3199736980, 0.0]],
                 [6, [2.2861906655, -1.3199736980, 0.0]], [6, [0.0, -2.6399473960, 0.0]],
                 [6, [-2.2861906655, -1.3199736980, 0.0]], [6, [-2.2861906655, 1.3199736980, 0.0]],
                 [1, [0.0, 4.6884105150, 0.0]], [1, [4.0602655512, 2.3442052575, 0.0]],
                 [1, [4.0602655512, -2.3442052575, 0.0]], [1, [0.0, -4.6884105150, 0.0]],
                 [1, [-4.0602655512, -2.3442052575, 0.0]], [1, [-4.0602655512, 2.3442052575, 0.0]]],
    }
    (OUT / "benchmark_geometries.json").write_text(json.dumps(geoms, indent=1))

    # H2O/6-31G* PySCF int2e values incl. Cartesian d (integrals/eri/mod.rs:4839-4925),
    # d-diagonal values already rescaled to per-component normalisation by the reference's test.
    h2o = {
        "geometry_bohr": [[8, [0.0, 0.0, 0.1173]], [1, [0.0, 0.7572, -0.4692]], [1, [0.0, -0.7572, -0.4692]]],
        "basis": "6-31g*", "tolerance": 5e-6,
        "values": [
            [[0, 0, 0, 0], 4.780445708111e+00], [[1, 1, 1, 1], 1.029677171970e+00],
            [[2, 2, 2, 2], 5.863293117932e-01], [[3, 3, 3, 3], 1.136875342624e+00],
            [[4, 4, 4, 4], 1.136875342624e+00], [[5, 5, 5, 5], 1.136875342624e+00],
            [[6, 6, 6, 6], 4.788356046311e-01], [[7, 7, 7, 7], 4.788356046311e-01],
            [[8, 8, 8, 8], 4.788356046311e-01], [[9, 9, 9, 9], 7.642154562066e-01],
            [[10, 10, 10, 10], 6.878539851698e-01], [[11, 11, 11, 11], 6.878539851698e-01],
            [[12, 12, 12, 12], 7.642154562066e-01], [[13, 13, 13, 13], 6.878539851698e-01],
            [[14, 14, 14, 14], 7.642154562066e-01], [[15, 15, 15, 15], 1.076566111405e+00],
            [[16, 16, 16, 16], 4.531503863486e-01], [[17, 17, 17, 17], 1.076566111405e+00],
            [[18, 18, 18, 18], 4.531503863486e-01], [[0, 0, 1, 1], 1.311436284253e+00],
            [[0, 0, 2, 2], 8.216638374470e-01], [[0, 1, 0, 1], 1.128540097925e-01],
            [[1, 2, 1, 2], 5.051255991433e-01], [[0, 0, 9, 9], 7.611442132861e-01],
            [[15, 15, 17, 17], 6.244949303271e-01],
        ],
    }
    (OUT / "h2o_631gs_eri.json").write_text(json.dumps(h2o, indent=1))

    # H2 STO-3G R=1.4 bohr PySCF J/K/G/F/D (scf/fock.rs:192-229), row-major 2x2.
    h2 = {
        "J": [1.345430416741971e+00, 8.933020749953189e-01, 8.933020749953189e-01, 1.345430416741970e+00],
        "K": [1.181117517432615e+00, 1.057614974304675e+00, 1.057614974304675e+00, 1.181117517432615e+00],
        "G": [7.548716580256631e-01, 3.644945878429816e-01, 3.644945878429816e-01, 7.548716580256630e-01],
        "F": [-3.655373508811572e-01, -5.938853765466354e-01, -5.938853765466354e-01, -3.655373508811574e-01],
        "D": [6.026571614189372e-01, 6.026571614189370e-01, 6.026571614189370e-01, 6.026571614189368e-01],
        "tolerance": 1e-10,
    }
    (OUT / "h2_fock.json").write_text(json.dumps(h2, indent=1))
    # libxc pointwise vectors quoted in functionals.rs tests are written by hand into
    # tests/golden/xc_libxc.json (numbers from functionals.rs:1753-1800, 2299-2320, 2427-2450, 2510-2533).
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
