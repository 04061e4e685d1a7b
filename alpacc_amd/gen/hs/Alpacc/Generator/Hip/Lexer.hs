-- | The lexer half of the HIP backend's table bundle (the lexer fields of @alpacc_tables@ in
-- include/alpacc_hip.h), computed from the byte-level lexer DFA that
-- "Alpacc.Generator.Analyzer" builds in mkLexer / mkLexerParser (lexerDFA over the
-- bytes of the token specification, src/Alpacc/Generator/Analyzer.hs:114, :204).
--
-- The reference's own backends receive the DFA already turned into encoded endomorphism ids
-- (IntParallelLexer, src/Alpacc/Lexer/Encode.hs:140-165). The HIP kernels need more than those
-- ids: the lexer's emit kernel walks the DFA itself from the state every tile starts in, so the
-- tables below include the DFA's step table and the state each endomorphism sends each state
-- to. They are built the same way as alpacc_amd/gen/endo.py and tables.py build them:
--
--   * a byte's endomorphism is the pair (state map, producing flag per source state) over the
--     DFA's states numbered 1..m in ascending order, 0 being the dead state
--     (src/Alpacc/Lexer/DFAParallelLexer.hs:39-68);
--   * "a then b" maps s to b(a(s)), and the move out of a(s) decides whether it produces
--     (DFAParallelLexer.hs:70-80);
--   * row 0 is the identity, row 1 the dead endomorphism (every state to 0), then the distinct
--     non-dead byte endomorphisms in byte order, then compositions in the order the closure finds
--     them: each round composes, in both orders, every pair of rows of which at least one was
--     added in the previous round, and a composition that sends every state to 0 is row 1;
--   * bytes with the same endomorphism share a byte class, numbered by first byte.
--
-- The emitted arrays (all row-major):
--
--   byte_class   u8[256]        class of each byte
--   endo_step    u16[e*C]       row of (endomorphism a, then one byte of class c)
--   endo_comp    u16[e*e]       row of (a then b), row a
--   endo_apply   u8[e*(m+1)]    state that a sends s to
--   dfa_step     u16[(m+1)*C]   next state | produce << 8 | terminal of next << 9
--   state_accept u8[m+1]        1 for an accepting state
module Alpacc.Generator.Hip.Lexer
  ( HipLexer (..),
    hipLexer,
  )
where

import Alpacc.Lexer.DFA (DFALexer, transitions')
import Alpacc.Lexer.FSA (FSA (..), Lexer (..))
import Control.Monad (when)
import Data.Array.Unboxed (UArray, (!))
import Data.Array.Unboxed qualified as UArray
import Data.Bits (shiftL, (.|.))
import Data.Containers.ListUtils (nubOrd)
import Data.Foldable (toList)
import Data.List (foldl')
import Data.Map.Strict (Map)
import Data.Map.Strict qualified as Map
import Data.Sequence (Seq, (|>))
import Data.Sequence qualified as Seq
import Data.Set qualified as Set
import Data.Text (Text)
import Data.Word (Word8)

data HipLexer = HipLexer
  { -- | e: endomorphism rows (identity 0, dead 1)
    numEndos :: !Int,
    -- | C: byte classes
    numClasses :: !Int,
    -- | m + 1: DFA states and the dead state 0
    numStates :: !Int,
    -- | the DFA's initial state, in 1..m
    initState :: !Int,
    byteClass :: [Int],
    endoStep :: [Int],
    endoComp :: [Int],
    endoApply :: [Int],
    dfaStep :: [Int],
    stateAccept :: [Int]
  }
  deriving (Show)

-- | State map and producing flag, indexed by source state 0..m.
data Endo = Endo !(UArray Int Int) !(UArray Int Bool)
  deriving (Eq, Ord, Show)

identityRow, deadRow :: Int
identityRow = 0
deadRow = 1

andThen :: Endo -> Endo -> Endo
andThen (Endo fa _) (Endo fb pb) =
  Endo (UArray.amap (fb !) fa) (UArray.amap (pb !) fa)

isDead :: Endo -> Bool
isDead (Endo f _) = all (== 0) (UArray.elems f)

type Rows = (Seq Endo, Map Endo Int)

-- | Append an endomorphism not seen before; the dead one is never appended (it is row 1).
addNew :: Rows -> Endo -> Rows
addNew acc@(rows, index) x
  | isDead x || Map.member x index = acc
  | otherwise = (rows |> x, Map.insert x (Seq.length rows) index)

-- | Close the rows under composition, round by round.
closeUnder :: Rows -> Either Text Rows
closeUnder = go 2
  where
    go done acc@(rows, _)
      | done >= e = Right acc
      | otherwise = bounded (foldl' addNew acc candidates) >>= go e
      where
        e = Seq.length rows
        row = Seq.index rows
        candidates =
          concat
            [ [andThen (row a) (row b) | b <- if a >= done then [2 .. e - 1] else [done .. e - 1]]
                ++ [andThen (row o) (row a) | a >= done, o <- [2 .. done - 1]]
            | a <- [2 .. e - 1]
            ]
    bounded acc@(rows, _)
      | Seq.length rows > 65536 =
          Left "Error: the lexer has more than 65536 endomorphisms, more than the HIP backend's 16-bit tables hold."
      | otherwise = Right acc

hipLexer ::
  (Ord s) =>
  -- | terminal id of a token label (the analyzer's terminal encoder)
  (k -> Maybe Integer) ->
  -- | terminal id written for "no token" (the analyzer's dead token)
  Integer ->
  DFALexer Word8 s k ->
  Either Text HipLexer
hipLexer terminalId deadTerminal lexer' = do
  when (width > 256) $
    Left "Error: the HIP backend's lexer takes at most 255 DFA states."
  when (any (\t -> t < 0 || t >= 128) (UArray.elems labelOf)) $
    Left "Error: the HIP backend's lexer takes terminal ids below 128."
  (rows, index) <- closeUnder start
  let e = Seq.length rows
      row = Seq.index rows
      compose a b
        | a == identityRow = b
        | b == identityRow = a
        | a == deadRow || b == deadRow = deadRow
        | otherwise =
            let c = andThen (row a) (row b)
             in if isDead c then deadRow else index Map.! c
      rowOf x = if isDead x then deadRow else index Map.! x
      byteRows = map rowOf byteEndos
      classRows = nubOrd byteRows
      classOf = Map.fromList (zip classRows [0 ..])
      step s c =
        let Endo f p = row c
            next = f ! s
         in next .|. (fromEnum (p ! s) `shiftL` 8) .|. ((labelOf ! next) `shiftL` 9)
  pure $
    HipLexer
      { numEndos = e,
        numClasses = length classRows,
        numStates = width,
        initState = number Map.! initial dfa,
        byteClass = map (classOf Map.!) byteRows,
        endoStep = [compose a c | a <- [0 .. e - 1], c <- classRows],
        endoComp = [compose a b | a <- [0 .. e - 1], b <- [0 .. e - 1]],
        endoApply = concat [UArray.elems f | Endo f _ <- toList rows],
        dfaStep = [step s c | s <- [0 .. width - 1], c <- classRows],
        stateAccept = 0 : [fromEnum (Set.member s (accepting dfa)) | s <- stateList]
      }
  where
    dfa = fsa lexer'
    stateList = Set.toAscList (states dfa)
    m = length stateList
    width = m + 1
    number = Map.fromList (zip stateList [1 ..])
    moves = transitions' dfa
    producing = producesToken lexer'
    labelOf :: UArray Int Int
    labelOf =
      UArray.listArray (0, m) $
        fromInteger deadTerminal
          : [ maybe (fromInteger deadTerminal) fromInteger (Map.lookup s (tokenMap lexer') >>= terminalId)
            | s <- stateList
            ]
    byteEndo b =
      Endo
        (UArray.listArray (0, m) (0 : [maybe 0 (number Map.!) (Map.lookup (s, b) moves) | s <- stateList]))
        (UArray.listArray (0, m) (False : [Map.member (s, b) moves && Set.member (s, b) producing | s <- stateList]))
    byteEndos = map byteEndo [minBound .. maxBound :: Word8]
    identityEndo =
      Endo (UArray.listArray (0, m) [0 .. m]) (UArray.listArray (0, m) (replicate width False))
    deadEndo =
      Endo (UArray.listArray (0, m) (replicate width 0)) (UArray.listArray (0, m) (replicate width False))
    start = foldl' addNew (Seq.fromList [identityEndo, deadEndo], Map.empty) byteEndos
