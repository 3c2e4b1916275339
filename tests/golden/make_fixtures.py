"""Writes the committed golden fixtures of this directory.

talib083_rma14.json: the 96-bar OHLC series and the TA-Lib 0.8.3 RMA(14) of
its close, taken from the reference's own fixture file
tests/fixtures/talib08_typed.json (produced there by TA-Lib 0.8.3 itself).
Only the data is copied: inputs and expected outputs.

Run from the repository root with the reference checkout at /root/reference.
"""
import json
import os

REF = "/root/reference/tests/fixtures/talib08_typed.json"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    d = json.load(open(REF))
    out = {"source": "reference tests/fixtures/talib08_typed.json (TA-Lib 0.8.3 output)",
           "function": "RMA(close, 14)",
           "open": d["open"], "high": d["high"], "low": d["low"], "close": d["close"],
           "rma14": d["rma"]}
    with open(os.path.join(HERE, "talib083_rma14.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
