"""python -m oracle: build the oracle library (make -C oracle)."""
import os
import subprocess

subprocess.run(["make", "-s", "-C", os.path.dirname(os.path.abspath(__file__))], check=True)
