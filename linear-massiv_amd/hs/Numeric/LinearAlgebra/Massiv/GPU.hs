{-# LANGUAGE AllowAmbiguousTypes #-}
{-# LANGUAGE DataKinds #-}
{-# LANGUAGE ForeignFunctionInterface #-}
{-# LANGUAGE ScopedTypeVariables #-}
{-# LANGUAGE TypeApplications #-}

-- |
-- Module      : Numeric.LinearAlgebra.Massiv.GPU
--
-- MI355X (gfx950) implementations of the dense-solve path of
-- linear-massiv, bound through the C ABI of @include/lmx.h@ (liblmx.so).
--
-- Every function keeps the signature of the CPU function it replaces
-- (same name, same 'Matrix' / 'Vector' types with type-level dimensions,
-- same argument order and meaning); only the module differs:
--
-- > import Numeric.LinearAlgebra.Massiv        hiding (matMulP, luSolveP, ...)
-- > import Numeric.LinearAlgebra.Massiv.GPU
--
-- Replaced functions (reference src/Numeric/LinearAlgebra/Massiv/...):
--
--   * 'matMulP'        BLAS/Level3.hs:183-207  (and matMul's RULE :296-300)
--   * 'gemm'           BLAS/Level3.hs:102-131
--   * 'lu'             Solve/LU.hs:121-167     (permutation vector output)
--   * 'luSolveP'       Solve/LU.hs:265-304;  'luSolve' Solve/LU.hs:240-259
--   * 'cholesky'       Solve/Cholesky.hs:114-147
--   * 'choleskySolveP' Solve/Cholesky.hs:241-273; 'choleskySolve' :229-235
--   * 'forwardSub', 'backSub', 'forwardSubUnit', 'backSubUnit'
--                      Solve/Triangular.hs:90-249
--
-- Error behaviour follows the reference: a singular or non-SPD input still
-- returns a result (Inf/NaN where the arithmetic produces them, LAPACK info
-- ignored); an infrastructure failure (no device, out of memory) throws
-- 'GpuError' -- there is no silent CPU fallback.
--
-- The ccalls are @safe@ (they run for milliseconds and block); arguments are
-- copied into pinned storable arrays first, which is O(n^2) against the
-- O(n^3) work.  This file is not built here (no GHC in this image); a
-- package that adds it lists liblmx in extra-libraries and include/ in
-- include-dirs.
module Numeric.LinearAlgebra.Massiv.GPU
  ( matMulP
  , gemm
  , lu
  , luFloat
  , luSolveP
  , luSolve
  , cholesky
  , choleskySolveP
  , choleskySolve
  , forwardSub
  , backSub
  , forwardSubUnit
  , backSubUnit
  , luSolveMixed
  , choleskySolveMixed
  , GpuError (..)
  ) where

import Control.Exception (Exception, throwIO)
import Control.Monad (when)
import Data.Int (Int32, Int64)
import qualified Data.Massiv.Array as M
import Foreign.C.String (CString, peekCString)
import Foreign.C.Types (CChar (..), CDouble (..), CInt (..))
import Foreign.Marshal.Alloc (allocaBytes)
import Foreign.Marshal.Array (allocaArray, peekArray)
import Foreign.Ptr (Ptr, castPtr)
import GHC.TypeNats (KnownNat)
import System.IO.Unsafe (unsafePerformIO)

import Numeric.LinearAlgebra.Massiv.Internal (dimVal)
import Numeric.LinearAlgebra.Massiv.Types (Matrix (..), Vector (..))

-- ------------------------------------------------------------------ FFI

foreign import ccall safe "lmx.h lmx_dgemm"
  c_dgemm :: CChar -> CChar -> Int64 -> Int64 -> Int64 -> CDouble -> Ptr Double -> Int64
          -> Ptr Double -> Int64 -> CDouble -> Ptr Double -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_dgetrf"
  c_dgetrf :: Int64 -> Ptr Double -> Int64 -> Ptr Int32 -> IO CInt
foreign import ccall safe "lmx.h lmx_sgetrf"
  c_sgetrf :: Int64 -> Ptr Float -> Int64 -> Ptr Int32 -> IO CInt
foreign import ccall safe "lmx.h lmx_dgesv"
  c_dgesv :: Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Int32 -> Ptr Double -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_dpotrf"
  c_dpotrf :: Int64 -> Ptr Double -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_dposv"
  c_dposv :: Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_dtrsm"
  c_dtrsm :: CChar -> CChar -> CChar -> Int64 -> Int64 -> Ptr Double -> Int64
          -> Ptr Double -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_dgesv_mixed"
  c_dgesv_mixed :: CInt -> Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> Int64
                -> Ptr Double -> Int64 -> Ptr () -> IO CInt
foreign import ccall safe "lmx.h lmx_dposv_mixed"
  c_dposv_mixed :: CInt -> Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> Int64
                -> Ptr Double -> Int64 -> Ptr () -> IO CInt
foreign import ccall unsafe "lmx.h lmx_last_error"
  c_last_error :: IO CString

-- | An infrastructure failure reported by liblmx (status < 0).
data GpuError = GpuError Int String deriving (Show)

instance Exception GpuError

-- status < 0 throws; info > 0 (singular / not positive definite) is
-- returned to the caller, which ignores it exactly as the reference does.
check :: CInt -> IO Int
check rc
  | rc < 0 = do
      msg <- c_last_error >>= peekCString
      throwIO (GpuError (fromIntegral rc) msg)
  | otherwise = pure (fromIntegral rc)

chr :: Char -> CChar
chr = toEnum . fromEnum

-- ------------------------------------------------------------------ marshalling

-- Row-major pinned copies: massiv 'M.S' arrays are pinned storable buffers.
toS2 :: M.Array M.P M.Ix2 Double -> M.Array M.S M.Ix2 Double
toS2 = M.computeAs M.S . M.delay

toS1 :: M.Array M.P M.Ix1 Double -> M.Array M.S M.Ix1 Double
toS1 = M.computeAs M.S . M.delay

fromS2 :: M.Array M.S M.Ix2 Double -> M.Array M.P M.Ix2 Double
fromS2 = M.computeAs M.P . M.delay

fromS1 :: M.Array M.S M.Ix1 Double -> M.Array M.P M.Ix1 Double
fromS1 = M.computeAs M.P . M.delay

-- | Runs an in-place device call on a fresh pinned copy and returns the copy.
inPlace2 :: M.Array M.P M.Ix2 Double -> (Ptr Double -> IO ()) -> IO (M.Array M.P M.Ix2 Double)
inPlace2 a f = do
  ma <- M.thawS (toS2 a)
  M.withPtr ma f
  fromS2 <$> M.freezeS ma

-- ------------------------------------------------------------------ GEMM

-- | C = A B (replaces matMulP, BLAS/Level3.hs:183-207).
matMulP :: forall m k n. (KnownNat m, KnownNat k, KnownNat n)
        => Matrix m k M.P Double -> Matrix k n M.P Double -> Matrix m n M.P Double
matMulP (MkMatrix a) (MkMatrix b) = unsafePerformIO $ do
  let mm = dimVal @m; kk = dimVal @k; nn = dimVal @n
      c0 = M.replicate M.Seq (M.Sz2 mm nn) 0 :: M.Array M.P M.Ix2 Double
  ms <- M.thawS (toS2 a)
  bs <- M.thawS (toS2 b)
  c <- inPlace2 c0 $ \pc -> M.withPtr ms $ \pa -> M.withPtr bs $ \pb ->
    () <$ (check =<< c_dgemm (chr 'N') (chr 'N') (fi mm) (fi nn) (fi kk) 1 pa (fi (max 1 kk)) pb
                             (fi (max 1 nn)) 0 pc (fi (max 1 nn)))
  pure (MkMatrix c)
{-# NOINLINE matMulP #-}

-- | alpha A B + beta C (replaces gemm, BLAS/Level3.hs:102-131).
gemm :: forall m k n. (KnownNat m, KnownNat k, KnownNat n)
     => Double -> Matrix m k M.P Double -> Matrix k n M.P Double -> Double
     -> Matrix m n M.P Double -> Matrix m n M.P Double
gemm alpha (MkMatrix a) (MkMatrix b) beta (MkMatrix c0) = unsafePerformIO $ do
  let mm = dimVal @m; kk = dimVal @k; nn = dimVal @n
  ms <- M.thawS (toS2 a)
  bs <- M.thawS (toS2 b)
  c <- inPlace2 c0 $ \pc -> M.withPtr ms $ \pa -> M.withPtr bs $ \pb ->
    () <$ (check =<< c_dgemm (chr 'N') (chr 'N') (fi mm) (fi nn) (fi kk) (realToFrac alpha) pa
                             (fi (max 1 kk)) pb (fi (max 1 nn)) (realToFrac beta) pc (fi (max 1 nn)))
  pure (MkMatrix c)
{-# NOINLINE gemm #-}

-- ------------------------------------------------------------------ LU

-- | Packed LU and the permutation vector (row i of PA is row piv!i of A),
-- the output format of lu (Solve/LU.hs:121-167).  The device returns the
-- LAPACK swap list; the permutation is its composition.
lu :: forall n. KnownNat n
   => Matrix n n M.P Double -> (Matrix n n M.P Double, M.Array M.P M.Ix1 Int)
lu (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  allocaArray (max 1 nn) $ \pipiv -> do
    f <- inPlace2 a $ \pa -> () <$ (check =<< c_dgetrf (fi nn) pa (fi (max 1 nn)) pipiv)
    swaps <- map fromIntegral <$> peekArray nn pipiv
    pure (MkMatrix f, M.fromList M.Seq (swapsToPerm nn swaps))
{-# NOINLINE lu #-}

-- | lu at e = Float (the reference's lu is generic over Fractional e,
-- Solve/LU.hs:121): the FP32 factorisation, with the same pivot rules.
luFloat :: forall n. KnownNat n
        => Matrix n n M.P Float -> (Matrix n n M.P Float, M.Array M.P M.Ix1 Int)
luFloat (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  allocaArray (max 1 nn) $ \pipiv -> do
    ma <- M.thawS (M.computeAs M.S (M.delay a) :: M.Array M.S M.Ix2 Float)
    M.withPtr ma $ \pa -> () <$ (check =<< c_sgetrf (fi nn) pa (fi (max 1 nn)) pipiv)
    f <- M.computeAs M.P . M.delay <$> M.freezeS ma
    swaps <- map fromIntegral <$> peekArray nn pipiv
    pure (MkMatrix f, M.fromList M.Seq (swapsToPerm nn swaps))
{-# NOINLINE luFloat #-}

swapsToPerm :: Int -> [Int] -> [Int]
swapsToPerm n swaps = M.toList (M.computeAs M.P (go (M.fromList M.Seq [0 .. n - 1]) (zip [0 ..] swaps)))
  where
    go :: M.Array M.P M.Ix1 Int -> [(Int, Int)] -> M.Array M.D M.Ix1 Int
    go p [] = M.delay p
    go p ((k, q) : rest)
      | k == q = go p rest
      | otherwise =
          let pk = p M.! k; pq = p M.! q
              p' = M.computeAs M.P (M.imap (\i x -> if i == k then pq else if i == q then pk else x) p)
          in go p' rest

-- | x with A x = b (replaces luSolveP, Solve/LU.hs:265-304).
luSolveP :: forall n. KnownNat n
         => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
luSolveP (MkMatrix a) (MkVector b) = unsafePerformIO $ do
  let nn = dimVal @n
  ma <- M.thawS (toS2 a)
  mb <- M.thawS (toS1 b)
  allocaArray (max 1 nn) $ \pipiv ->
    M.withPtr ma $ \pa -> M.withPtr mb $ \pb ->
      () <$ (check =<< c_dgesv (fi nn) 1 pa (fi (max 1 nn)) pipiv pb 1)
  MkVector . fromS1 <$> M.freezeS mb
{-# NOINLINE luSolveP #-}

-- | The generic luSolve (Solve/LU.hs:240-259) at P Double.
luSolve :: forall n. KnownNat n
        => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
luSolve = luSolveP

-- ------------------------------------------------------------------ Cholesky

-- | Lower factor G with A = G G^T, upper triangle zero
-- (replaces cholesky, Solve/Cholesky.hs:114-147).
cholesky :: forall n. KnownNat n => Matrix n n M.P Double -> Matrix n n M.P Double
cholesky (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  g <- inPlace2 a $ \pa -> () <$ (check =<< c_dpotrf (fi nn) pa (fi (max 1 nn)))
  pure (MkMatrix (M.computeAs M.P (M.imap (\(i M.:. j) x -> if j > i then 0 else x) g)))
{-# NOINLINE cholesky #-}

-- | Replaces choleskySolveP (Solve/Cholesky.hs:241-273).
choleskySolveP :: forall n. KnownNat n
               => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
choleskySolveP (MkMatrix a) (MkVector b) = unsafePerformIO $ do
  let nn = dimVal @n
  ma <- M.thawS (toS2 a)
  mb <- M.thawS (toS1 b)
  M.withPtr ma $ \pa -> M.withPtr mb $ \pb ->
    () <$ (check =<< c_dposv (fi nn) 1 pa (fi (max 1 nn)) pb 1)
  MkVector . fromS1 <$> M.freezeS mb
{-# NOINLINE choleskySolveP #-}

choleskySolve :: forall n. KnownNat n
              => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
choleskySolve = choleskySolveP

-- ------------------------------------------------------------------ triangular

trsv :: forall n. KnownNat n => Char -> Char
     -> Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
trsv uplo diag (MkMatrix t) (MkVector b) = unsafePerformIO $ do
  let nn = dimVal @n
  mt <- M.thawS (toS2 t)
  mb <- M.thawS (toS1 b)
  M.withPtr mt $ \pt -> M.withPtr mb $ \pb ->
    () <$ (check =<< c_dtrsm (chr uplo) (chr 'N') (chr diag) (fi nn) 1 pt (fi (max 1 nn)) pb 1)
  MkVector . fromS1 <$> M.freezeS mb
{-# NOINLINE trsv #-}

-- | Solve/Triangular.hs:90-107
forwardSub :: forall n. KnownNat n => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
forwardSub = trsv @n 'L' 'N'

-- | Solve/Triangular.hs:136-153
backSub :: forall n. KnownNat n => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
backSub = trsv @n 'U' 'N'

-- | Solve/Triangular.hs:187-201
forwardSubUnit :: forall n. KnownNat n => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
forwardSubUnit = trsv @n 'L' 'U'

-- | Solve/Triangular.hs:235-249
backSubUnit :: forall n. KnownNat n => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
backSubUnit = trsv @n 'U' 'U'

-- ------------------------------------------------------------------ mixed precision (new surface)

-- | FP64-accurate solve with an FP32 factorisation and FP64 refinement on
-- the device; falls back to the FP64 factorisation when refinement does
-- not converge.  No reference counterpart.
luSolveMixed :: forall n. KnownNat n
             => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
luSolveMixed = mixed @n c_dgesv_mixed

choleskySolveMixed :: forall n. KnownNat n
                   => Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
choleskySolveMixed = mixed @n c_dposv_mixed

mixed :: forall n. KnownNat n
      => (CInt -> Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> Int64
          -> Ptr Double -> Int64 -> Ptr () -> IO CInt)
      -> Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
mixed call (MkMatrix a) (MkVector b) = unsafePerformIO $ do
  let nn = dimVal @n
  ma <- M.thawS (toS2 a)
  mb <- M.thawS (toS1 b)
  mx <- M.thawS (toS1 b)
  allocaBytes 32 $ \info ->
    M.withPtr ma $ \pa -> M.withPtr mb $ \pb -> M.withPtr mx $ \px -> do
      rc <- call 1 (fi nn) 1 pa (fi (max 1 nn)) pb 1 px 1 (castPtr info)
      when (rc < 0) (() <$ check rc)
  MkVector . fromS1 <$> M.freezeS mx
{-# NOINLINE mixed #-}

fi :: Int -> Int64
fi = fromIntegral
