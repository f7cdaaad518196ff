"""Imports the product package (directory name spooky-git_amd) for the tests."""
import importlib
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

spk = importlib.import_module("spooky-git_amd")
