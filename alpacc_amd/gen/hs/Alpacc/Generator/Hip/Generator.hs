-- | The @hip@ backend: the analyzer's tables as a C header for the MI355X library's C ABI
-- (include/alpacc_hip.h). It takes the slot of the CUDA backend
-- (src/Alpacc/Generator/Cuda/Generator.hs:25-82): where that one embeds a CUDA program around the
-- tables, this one writes only the tables, @static const@ arrays and one @alpacc_tables@ value
-- named @<name>_tables@, plus the terminal and production names (Analyzer.hs:97-147) that the
-- reference's @lex@ / @parse@ entries return. A C or C++ program includes the header and calls
-- @alpacc_create(&<name>_tables, device)@.
--
-- The layout is the one alpacc_amd/gen/emit_c.py writes, array for array and field for field;
-- the parser arrays are equal to that generator's, the lexer arrays describe the same lexer
-- (see "Alpacc.Generator.Hip.Lexer").
module Alpacc.Generator.Hip.Generator
  ( generator,
    namedGenerator,
    hipHeader,
  )
where

import Alpacc.Generator.Analyzer
import Alpacc.Generator.Hip.Lexer
import Alpacc.Generator.Hip.Parser
import Data.ByteString qualified as ByteString
import Data.Char (chr, isAlphaNum, isAscii, isDigit, toUpper)
import Data.List.Split (chunksOf)
import Data.Maybe (fromMaybe)
import Data.Text (Text)
import Data.Text qualified as Text
import Data.Text.Encoding qualified as Text
import Numeric (showOct)

-- | Identifiers prefixed @parser@, as the CLI names the output of standard input.
generator :: Generator [Text]
generator = namedGenerator "parser"

-- | Identifiers prefixed with a name (the CLI would pass the grammar file's base name).
namedGenerator :: Text -> Generator [Text]
namedGenerator name =
  Generator {generate = either (error . Text.unpack) id . hipHeader name}

tshow :: (Show a) => a -> Text
tshow = Text.pack . show

identifier :: Text -> Text
identifier name
  | Text.null s || isDigit (Text.head s) = "g_" <> s
  | otherwise = s
  where
    s = Text.map (\c -> if isAscii c && (isAlphaNum c || c == '_') then c else '_') name

-- | @static const ctype name[n] = { ... };@ with perLine values a line.
cArray :: Text -> Text -> Int -> Text -> [Integer] -> Text
cArray ctype name perLine suffix values =
  Text.concat
    [ "static const ",
      ctype,
      " ",
      name,
      "[",
      tshow (length vals),
      "] = {\n  ",
      Text.intercalate ",\n  " (Text.intercalate ", " <$> chunksOf perLine vals),
      "\n};\n"
    ]
  where
    vals = if null values then ["0"] else (<> suffix) . tshow <$> values

-- | A C string literal of the text's UTF-8 bytes.
cString :: Text -> Text
cString s = "\"" <> Text.concat (escape <$> ByteString.unpack (Text.encodeUtf8 s)) <> "\""
  where
    escape b
      | b == 0x22 || b == 0x5C = Text.pack ['\\', chr (fromIntegral b)]
      | b >= 0x20 && b < 0x7F = Text.singleton (chr (fromIntegral b))
      | otherwise = Text.pack ('\\' : pad (showOct b ""))
    pad x = replicate (3 - length x) '0' ++ x

cNames :: Text -> Text -> [Text] -> [Text]
cNames p what names =
  [ "#define " <> Text.toUpper p <> "_NUM_" <> what <> "_NAMES " <> tshow (length names),
    "static const char *const "
      <> p
      <> "_"
      <> Text.toLower what
      <> "_names["
      <> tshow (max 1 (length names))
      <> "] = {\n  "
      <> Text.intercalate ",\n  " (cString <$> if null names then [""] else names)
      <> "\n};"
  ]

lexerPart :: Text -> Lexer -> Either Text ([Text], [(Text, Text)])
lexerPart p lx = do
  t <- hipLexer Just (deadToken lx) (byteDFA lx)
  let arr ctype field values = (cArray ctype (p <> "_" <> field) 16 "" (toInteger <$> values), (field, p <> "_" <> field))
      arrays =
        [ arr "uint8_t" "byte_class" (byteClass t),
          arr "uint16_t" "endo_step" (endoStep t),
          arr "uint16_t" "endo_comp" (endoComp t),
          arr "uint8_t" "endo_apply" (endoApply t),
          arr "uint16_t" "dfa_step" (dfaStep t),
          arr "uint8_t" "state_accept" (stateAccept t)
        ]
      scalars =
        [ ("num_endos", tshow (numEndos t)),
          ("num_classes", tshow (numClasses t)),
          ("num_states", tshow (numStates t)),
          ("init_state", tshow (initState t)),
          ("ignore_terminal", tshow (fromMaybe (-1) (ignoreToken lx))),
          ("dead_terminal", tshow (deadToken lx))
        ]
  pure (fst <$> arrays, scalars ++ (snd <$> arrays))

parserPart :: Text -> Parser -> Either Text ([Text], [(Text, Text)])
parserPart p parser = do
  t <- hipParser parser
  let name field = p <> "_" <> field
      hashed = not (null (pairHashKeys t))
      dense = not (null (pairIndex t))
      arrays =
        [ cArray "int32_t" (name "arities") 16 "" (productionArity t),
          cArray "int32_t" (name "prod_to_terminal") 16 "" (productionTerminal t),
          cArray "uint32_t" (name "pair_index") 16 "" (pairIndex t)
        ]
          ++ ( if hashed
                 then
                   [ cArray "uint64_t" (name "pair_hash_keys") 8 "ull" (toInteger <$> pairHashKeys t),
                     cArray "uint32_t" (name "pair_hash_vals") 16 "" (pairHashVals t)
                   ]
                 else []
             )
          ++ [ cArray "uint32_t" (name "pair_entry") 16 "" (toInteger <$> pairEntry t),
               cArray "uint16_t" (name "stacks") 16 "" (stacks t),
               cArray "uint16_t" (name "productions") 16 "" (toInteger <$> productions t)
             ]
      fields =
        [ ("q", tshow (lookback parser)),
          ("k", tshow (lookahead parser)),
          ("num_terminals", tshow (numTerminals t)),
          ("start_terminal", tshow (startTerminal parser)),
          ("end_terminal", tshow (endTerminal parser)),
          ("empty_terminal", tshow (emptyTerminal parser)),
          ("num_productions", tshow (numberOfProductions parser)),
          ("arities", name "arities"),
          ("prod_to_terminal", name "prod_to_terminal"),
          ("pair_index_len", tshow (length (pairIndex t)) <> "ull"),
          ("pair_index", if dense then name "pair_index" else "0"),
          ("num_pair_entries", tshow (length (pairEntry t) `div` 4)),
          ("pair_entry", name "pair_entry"),
          ("stacks_len", tshow (length (stacks t))),
          ("stacks", name "stacks"),
          ("productions_len", tshow (length (productions t))),
          ("productions", name "productions")
        ]
          ++ ( if hashed
                 then
                   [ ("pair_hash_len", tshow (length (pairHashKeys t)) <> "ull"),
                     ("pair_hash_keys", name "pair_hash_keys"),
                     ("pair_hash_vals", name "pair_hash_vals")
                   ]
                 else []
             )
  pure (arrays, fields)

-- | The header text, or why the grammar does not fit the HIP backend's tables.
hipHeader :: Text -> Analyzer [Text] -> Either Text Text
hipHeader name analyzer = do
  (lexArrays, lexFields) <- maybe (Right ([], [])) (lexerPart p) maybeLexer
  (parArrays, parFields) <- maybe (Right ([], [])) (parserPart p) maybeParser
  pure . Text.intercalate "\n" $
    concat
      [ ["/* Tables of the '" <> name <> "' grammar (" <> mode <> "), generated by alpacc hip. */"],
        ("// " <>) <$> meta analyzer,
        ["#ifndef " <> guard, "#define " <> guard, "", "#include \"alpacc_hip.h\"", ""],
        lexArrays,
        parArrays,
        ["static const alpacc_tables " <> p <> "_tables = {"],
        (\(k, v) -> "  ." <> k <> " = " <> v <> ",") <$> lexFields ++ parFields,
        ["};\n"],
        cNames p "TERMINAL" (terminalToName analyzer),
        cNames p "PRODUCTION" (maybe [] productionToName maybeParser),
        ["", "#endif /* " <> guard <> " */", ""]
      ]
  where
    p = identifier name
    guard = "ALPACC_TABLES_" <> Text.map toUpper p <> "_H"
    (maybeLexer, maybeParser, mode) =
      case analyzerKind analyzer of
        Lex l -> (Just l, Nothing, "lexer")
        Parse q -> (Nothing, Just q, "parser")
        Both l q -> (Just l, Just q, "both")
