"""Three inputs of kolmo's determinism check, regenerated byte for byte so that whole frames can be compared with the
digests kolmo publishes for them (benchmark/determinism/expected.sha256). Test infrastructure only.

The bytes must equal kolmo's, so the draws follow benchmark/determinism/check.py:47-52 in the same order from
Python's Mersenne Twister seeded with 20261116:
  text.txt   110,000 words from a 3,000-word vocabulary ("w<i>" repeated 1 + i % 3 times), word index
             min(int(paretovariate(1.1)) - 1, 2999), a newline after a word with probability 0.08, joined by spaces;
  table.csv  a header and 20,000 rows "C<i//40>,<1980 + i%40>,<1000000 + 37 i>,<randint(0, 99999)/100 :.2f>,
             <random() :.4f>", continuing the same generator;
  log.txt    15,000 lines "2026-11-16T10:<(i//60)%60 :02d>:<i%60 :02d>Z host<i%7> GET /api/item/<randint(1, 5000)> 200
             <randint(1, 900)>ms", continuing the same generator (check.py:51).
"""
import random


def determinism_inputs():
    """[(name, bytes)] for text.txt, table.csv and log.txt."""
    rng = random.Random(20261116)
    vocab = [f"w{i}" * (1 + i % 3) for i in range(3000)]
    parts = []
    for _ in range(110000):
        word = vocab[min(int(rng.paretovariate(1.1)) - 1, 2999)]
        parts.append(word + ("\n" if rng.random() < 0.08 else ""))
    text = " ".join(parts).encode()
    lines = ["country,year,population,co2,share"]
    for i in range(20000):
        co2 = rng.randint(0, 99999) / 100
        share = rng.random()
        lines.append(f"C{i // 40},{1980 + i % 40},{1000000 + i * 37},{co2:.2f},{share:.4f}")
    table = ("\n".join(lines) + "\n").encode()
    log = []
    for i in range(15000):
        item = rng.randint(1, 5000)
        ms = rng.randint(1, 900)
        log.append(f"2026-11-16T10:{(i // 60) % 60:02d}:{i % 60:02d}Z host{i % 7} GET /api/item/{item} 200 {ms}ms\n")
    return [("text.txt", text), ("table.csv", table), ("log.txt", "".join(log).encode())]
