-- | The parser half of the HIP backend's table bundle (the parser fields of @alpacc_tables@ in
-- include/alpacc_hip.h), computed from the analyzer's LLP table
-- (src/Alpacc/Generator/Analyzer.hs:52-66; built by llpHashTable, src/Alpacc/Encode.hs:353-375).
--
-- The reference's table is one open-addressing array over the padded key windows (the q
-- terminals before a position and the k from it on), each key holding a span of the flattened
-- bracket stacks and a span of the flattened productions. The HIP kernels look a window up by
-- its dense index (the terminals as digits base num_terminals, the first most significant), and
-- give every distinct (brackets, productions) pair one entry:
--
--   pair_entry   u32[E*4]      stack start, stack end, production start, production end
--   stacks       u16[]         brackets: symbol | 0x8000 for a left bracket
--   productions  u16[]         production ids
--   pair_index   u32[NT^(q+k)] entry of every window, 0xFFFFFFFF for none (up to 2^26 windows)
--   pair_hash_keys u64[H], pair_hash_vals u32[H]: beyond 2^26 windows, the windows' dense
--                indices hashed instead: home slot (index * 0x9E3779B97F4A7C15 mod 2^64) >>
--                (64 - log2 H), linear probing, H the smallest power of two from 16 up that is
--                at least twice the number of windows; free slots hold 0xFFFFFFFFFFFFFFFF
--
-- Entries are numbered in ascending key order (the order of the analyzer's Map), as
-- alpacc_amd/gen/tables.py numbers them, so both generators emit the same parser arrays.
module Alpacc.Generator.Hip.Parser
  ( HipParser (..),
    hipParser,
  )
where

import Alpacc.Encode (LLPTable (..))
import Alpacc.Generator.Analyzer (Parser (..))
import Alpacc.HashTable (OpenAddressing (..))
import Alpacc.Types (IntType (..))
import Control.Monad (when)
import Data.Array qualified as Array
import Data.Bits (clearBit, countTrailingZeros, shiftR, testBit, (.&.), (.|.))
import Data.IntMap.Strict qualified as IntMap
import Data.List (foldl')
import Data.Map.Strict qualified as Map
import Data.Text (Text)
import Data.Word (Word64)

data HipParser = HipParser
  { -- | terminal ids are [0, numTerminals): the grammar's terminals, ignore, empty, start, end
    numTerminals :: !Integer,
    -- | dense pair table (empty when hashed)
    pairIndex :: [Integer],
    pairHashKeys :: [Word64],
    pairHashVals :: [Integer],
    pairEntry :: [Int],
    stacks :: [Integer],
    productions :: [Int],
    -- | arity of each production
    productionArity :: [Integer],
    -- | terminal of each terminal production, -1 for the others
    productionTerminal :: [Integer]
  }
  deriving (Show)

missing :: Integer
missing = 0xFFFFFFFF

leftBit :: Integer
leftBit = 0x8000

denseMax :: Integer
denseMax = 2 ^ (26 :: Int)

hashEmpty :: Word64
hashEmpty = maxBound

hashMul :: Word64
hashMul = 0x9E3779B97F4A7C15

slice :: Int -> Int -> [a] -> [a]
slice from to = take (to - from) . drop from

hipParser :: Parser -> Either Text HipParser
hipParser parser = do
  when (any ((>= leftBit) . snd . decode) flatStacks) $
    Left "Error: the HIP backend's brackets hold symbol ids below 2^15."
  when (hashed && windows >= toInteger hashEmpty) $
    Left "Error: the key windows do not fit 64 bits."
  pure $
    HipParser
      { numTerminals = nT,
        pairIndex =
          if hashed
            then []
            else
              let byIndex = Map.fromList keyEntry
               in [Map.findWithDefault missing i byIndex | i <- [0 .. windows - 1]],
        pairHashKeys = if hashed then map fst hashSlots else [],
        pairHashVals = if hashed then map snd hashSlots else [],
        pairEntry = concat (reverse entriesRev),
        stacks = concat (reverse stacksRev),
        productions = concat (reverse prodsRev),
        productionArity = arities parser,
        productionTerminal = maybe (-1) id <$> productionToTerminal parser
      }
  where
    nT = 1 + max (startTerminal parser) (endTerminal parser)
    w = lookback parser + lookahead parser
    windows = nT ^ w
    hashed = windows > denseMax
    LLPTable flatStacks flatProds oa = llpTable parser
    nbs = fromInteger (numBits (bracketType parser)) :: Int
    decode b
      | testBit b (nbs - 1) = (True, clearBit b (nbs - 1))
      | otherwise = (False, b)
    device (left, s) = if left then s .|. leftBit else s
    -- every key with its (brackets, productions), in ascending key order
    keyed =
      Map.toAscList $
        Map.fromList
          [ (key, (map (device . decode) (slice s0 s1 flatStacks), slice p0 p1 flatProds))
          | (True, key, ((s0, s1), (p0, p1))) <- Array.elems (oaArray oa)
          ]
    denseIndex = foldl' (\acc t -> acc * nT + t) 0
    -- one entry per distinct (brackets, productions), numbered by first key
    (entriesRev, stacksRev, prodsRev, keyEntryRev, _, _, _) =
      foldl' addKey ([], [], [], [], Map.empty, 0 :: Int, 0 :: Int) keyed
    addKey (es, ss, ps, kes, seen, sLen, pLen) (key, sig@(br, pr)) =
      case Map.lookup sig seen of
        Just entry -> (es, ss, ps, (denseIndex key, entry) : kes, seen, sLen, pLen)
        Nothing ->
          let entry = toInteger (Map.size seen)
              sLen' = sLen + length br
              pLen' = pLen + length pr
           in ( [sLen, sLen', pLen, pLen'] : es,
                br : ss,
                pr : ps,
                (denseIndex key, entry) : kes,
                Map.insert sig entry seen,
                sLen',
                pLen'
              )
    keyEntry = reverse keyEntryRev
    -- the hashed table: keys inserted in ascending key order
    hashSize = head [s | s <- iterate (* 2) (16 :: Int), s >= 2 * length keyEntry]
    hashBits = countTrailingZeros hashSize
    home i = fromIntegral ((fromInteger i * hashMul) `shiftR` (64 - hashBits)) :: Int
    place slots (i, entry) = IntMap.insert (probe (home i)) (fromInteger i, entry) slots
      where
        probe h
          | IntMap.member h slots = probe ((h + 1) .&. (hashSize - 1))
          | otherwise = h
    hashSlots =
      IntMap.elems $
        IntMap.union
          (foldl' place IntMap.empty keyEntry)
          (IntMap.fromList [(h, (hashEmpty, missing)) | h <- [0 .. hashSize - 1]])
