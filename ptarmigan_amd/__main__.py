"""Command line: `python -m ptarmigan_amd input.yml`, as the reference's
`ptarmigan input.yml` (src/main.rs:321-335, 1725-1757): reads the deck, runs
every a0 value on GPU 0, and writes the requested distributions and stats.txt
next to the input file. Exits non-zero with the error on a bad deck."""
import os
import sys


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m ptarmigan_amd <input.yml>", file=sys.stderr)
        return 2
    from .deck import Deck, DeckError, run_deck
    from .engine import PtgError
    path = argv[0]
    try:
        deck = Deck.from_file(path)
        res = run_deck(deck, out_dir=os.path.dirname(os.path.abspath(path)))
    except (OSError, DeckError, PtgError) as err:
        print(f"error: {err}", file=sys.stderr)
        return 1
    for a0, part, cnt in res:
        n = {k: int((part["species"] == s).sum()) for k, s in (("electrons", 0), ("photons", 2), ("positrons", 1))}
        print(f"a0 = {a0:g}: {n['electrons']} electrons, {n['photons']} photons, {n['positrons']} positrons in the "
              f"final state ({cnt['lepton_steps']} lepton steps, {cnt['photon_steps']} photon steps)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
