"""CPU oracle for the alpacc lexer/parser hot path: test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use this package,
always as the checker or the reported CPU baseline, never as the thing measured or shipped.

Contents:
  * ``lexer_fut``  - literal restatement of futhark/lexer.fut (scan over encoded endomorphisms,
                     including the 16 MiB chunk loop of ``lex_int_flag``);
  * ``parser_fut`` - literal restatement of futhark/parser.fut (hash lookup, segmented copy,
                     depth scan, min-tree ``find_previous``, parents, terminal nodes);
  * ``sequential`` - independent sequential oracles: ``tokenize`` (src/Alpacc/Lexer/DFA.hs:415-447)
                     and ``llpParse`` + tree flattening (src/Alpacc/LLP.hs:600-745);
  * ``wire``       - the test.fut / Data.Binary wire formats (big-endian u64 records);
  * ``c/``         - a C/OpenMP restatement of the same data-parallel algorithm, for large inputs
                     and the CPU baseline.
"""
