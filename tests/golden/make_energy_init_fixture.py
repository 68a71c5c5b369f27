"""Writes energy_init.json: the reference's known-answer vectors for the ATTRACT pair
potential (energy-init/make-energy-init.out distances, energy-init/score.out ATTRACT
FP64 energies, parameters rbc=3.88, abc=13.81, ipon=1 of energy-init/energy-init.py:1-3).
Run in a checkout that has the reference mounted at /root/reference."""
import json
import os

REF = "/root/reference/energy-init"
HERE = os.path.dirname(os.path.abspath(__file__))

dist = [float(l) for l in open(os.path.join(REF, "make-energy-init.out")) if l.strip()]
ene = [float(l.split()[1]) for l in open(os.path.join(REF, "score.out")) if l.strip()]
assert len(dist) == len(ene)
json.dump({"rbc": 3.88, "abc": 13.81, "ipon": 1, "potshape": 8, "distance": dist, "attract_energy": ene},
          open(os.path.join(HERE, "energy_init.json"), "w"), indent=0)
print(len(dist), "points")
