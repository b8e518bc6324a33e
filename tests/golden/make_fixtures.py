"""Copy the reference's data fixtures (golden outputs into tests/golden, mechanism inputs
into the package's mechanisms/ folder, from which the product reads them).

Run once in a container where the reference checkout is mounted; the copied files are
plain data (Cantera YAML mechanism inputs and whitespace-separated result tables) and are
committed, so nothing at test time reads the reference checkout.

Sources (paths relative to the reference root; mechanisms/ is prop-streams2_amd/mechanisms):
  examples/laminar_premix_flame/input_cantera.yaml   -> mechanisms/h2_premix_11sp.yaml
  examples/reactive_shock_tube/input_cantera.yaml    -> mechanisms/h2_shocktube_9sp.yaml
  examples/batch_reactor/input_cantera.yaml          -> mechanisms/keromnes2013_14sp.yaml
  examples/syngas_mech_eval/Davis2005.yaml           -> mechanisms/davis2005_13sp.yaml
  examples/syngas_mech_eval/Varga2016.yaml           -> mechanisms/varga2016_14sp.yaml
  examples/batch_reactor/OUTPUT/cantera_results.dat  -> batch_reactor_cantera.dat
  examples/syngas_mech_eval/OUTPUT/output_*.dat      -> syngas_<mech>_streams.dat
  examples/reactive_shock_tube/OUTPUT/Ferrer_*.dat   -> shocktube/Ferrer_*.dat
  examples/batch_reactor/multideal.ini               -> batch_reactor.ini
  examples/reactive_shock_tube/multideal.ini         -> reactive_shock_tube.ini
"""
import pathlib
import shutil
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = pathlib.Path(__file__).resolve().parent
MECH_OUT = OUT.parents[1] / "prop-streams2_amd"

COPIES = {
    "examples/laminar_premix_flame/input_cantera.yaml": "mechanisms/h2_premix_11sp.yaml",
    "examples/reactive_shock_tube/input_cantera.yaml": "mechanisms/h2_shocktube_9sp.yaml",
    "examples/batch_reactor/input_cantera.yaml": "mechanisms/keromnes2013_14sp.yaml",
    "examples/syngas_mech_eval/Davis2005.yaml": "mechanisms/davis2005_13sp.yaml",
    "examples/syngas_mech_eval/Varga2016.yaml": "mechanisms/varga2016_14sp.yaml",
    "examples/batch_reactor/OUTPUT/cantera_results.dat": "batch_reactor_cantera.dat",
    "examples/syngas_mech_eval/OUTPUT/output_Keromnes2013.dat": "syngas_keromnes2013_streams.dat",
    "examples/syngas_mech_eval/OUTPUT/output_Davis2005.dat": "syngas_davis2005_streams.dat",
    "examples/syngas_mech_eval/OUTPUT/output_Varga2016.dat": "syngas_varga2016_streams.dat",
    "examples/batch_reactor/multideal.ini": "batch_reactor.ini",
    "examples/reactive_shock_tube/multideal.ini": "reactive_shock_tube.ini",
}


def main():
    for src, dst in COPIES.items():
        d = (MECH_OUT if dst.startswith("mechanisms/") else OUT) / dst
        d.parent.mkdir(parents=True, exist_ok=True)
        data = (REF / src).read_bytes().replace(b"\r\n", b"\n")
        d.write_bytes(data)
    st = REF / "examples/reactive_shock_tube/OUTPUT"
    (OUT / "shocktube").mkdir(exist_ok=True)
    for f in sorted(st.glob("Ferrer_*.dat")):
        shutil.copyfile(f, OUT / "shocktube" / f.name)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
