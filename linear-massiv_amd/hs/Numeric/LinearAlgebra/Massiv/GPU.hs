{-# LANGUAGE AllowAmbiguousTypes #-}
{-# LANGUAGE DataKinds #-}
{-# LANGUAGE KindSignatures #-}
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
--   * 'det'            Solve/LU.hs:534-558
--   * 'inverseIteration' Eigen/Power.hs:113-135
--   * 'matvecP', 'matvec', 'gemv'  BLAS/Level2.hs:92-144
--   * 'norm1', 'normInf', 'normFrob', 'vnorm1', 'vnorm2', 'vnormInf'
--                      Norms.hs:72-148
--   * 'matMulAtAP'     BLAS/Level3.hs:353-373
--
-- New surface (no reference counterpart): resident factors ('luFactor',
-- 'luSolveWith', 'luSolveManyWith', 'choleskyFactor', 'choleskySolveWith',
-- at FP64 or a refined FP32 / BF16 precision), the refined solvers
-- ('luSolveMixed', 'choleskySolveMixed', 'luSolveMixedWith',
-- 'choleskySolveMixedWith') and the BF16 matrix product ('matMulBF16').
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
  , luSolveMixedWith
  , choleskySolveMixedWith
  , Precision (..)
  , det
  , inverseIteration
  , matvecP
  , matvec
  , gemv
  , norm1
  , normInf
  , normFrob
  , vnorm1
  , vnorm2
  , vnormInf
  , matMulAtAP
  , matMulBF16
  , Factor
  , luFactor
  , choleskyFactor
  , luSolveWith
  , luSolveManyWith
  , choleskySolveWith
  , GpuError (..)
  ) where

import Control.Exception (Exception, throwIO)
import Control.Monad (void, when)
import Data.Int (Int32, Int64)
import Data.Word (Word16)
import Foreign.Concurrent (newForeignPtr)
import Foreign.ForeignPtr (ForeignPtr, withForeignPtr)
import Foreign.Marshal.Alloc (alloca)
import Foreign.Storable (peek)
import qualified Data.Massiv.Array as M
import Foreign.C.String (CString, peekCString)
import Foreign.C.Types (CChar (..), CDouble (..), CInt (..))
import Foreign.Marshal.Alloc (allocaBytes)
import Foreign.Marshal.Array (allocaArray, peekArray)
import Foreign.Ptr (Ptr, castPtr)
import GHC.TypeNats (KnownNat, Nat)
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
foreign import ccall safe "lmx.h lmx_dlu"
  c_dlu :: Int64 -> Ptr Double -> Int64 -> Ptr Int32 -> IO CInt
foreign import ccall safe "lmx.h lmx_slu"
  c_slu :: Int64 -> Ptr Float -> Int64 -> Ptr Int32 -> IO CInt
foreign import ccall safe "lmx.h lmx_ddet"
  c_ddet :: Int64 -> Ptr Double -> Int64 -> Ptr Double -> IO CInt
foreign import ccall safe "lmx.h lmx_dinverse_iteration"
  c_dinverse_iteration :: Int64 -> Ptr Double -> Int64 -> CDouble -> Ptr Double -> Int32 -> CDouble
                       -> Ptr Double -> Ptr Double -> Ptr Int32 -> IO CInt
foreign import ccall safe "lmx.h lmx_dgemv"
  c_dgemv :: Int64 -> Int64 -> CDouble -> Ptr Double -> Int64 -> Ptr Double -> CDouble -> Ptr Double
          -> IO CInt
foreign import ccall safe "lmx.h lmx_dnorm"
  c_dnorm :: CChar -> Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> IO CInt
foreign import ccall safe "lmx.h lmx_dsyrk"
  c_dsyrk :: CChar -> Int64 -> Int64 -> CDouble -> Ptr Double -> Int64 -> CDouble -> Ptr Double -> Int64
          -> IO CInt
foreign import ccall safe "lmx.h lmx_gemm_bf16"
  c_gemm_bf16 :: CChar -> CChar -> Int64 -> Int64 -> Int64 -> Float -> Ptr Word16 -> Int64 -> Ptr Word16
              -> Int64 -> Float -> Ptr Float -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_factor_create"
  c_factor_create :: Ptr (Ptr ()) -> CInt -> CInt -> Int64 -> Ptr Double -> Int64 -> IO CInt
foreign import ccall safe "lmx.h lmx_factor_solve"
  c_factor_solve :: Ptr () -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> Int64 -> Ptr () -> IO CInt
foreign import ccall safe "lmx.h lmx_factor_release"
  c_factor_release :: Ptr () -> IO CInt
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
-- the output of lu (Solve/LU.hs:121-167), with lu's own pivot rule: the
-- last row of largest |a_ik| wins a tie (maximumBy, :128-135).  lmx_dlu
-- returns the permutation vector itself.
lu :: forall n. KnownNat n
   => Matrix n n M.P Double -> (Matrix n n M.P Double, M.Array M.P M.Ix1 Int)
lu (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  allocaArray (max 1 nn) $ \pperm -> do
    f <- inPlace2 a $ \pa -> () <$ (check =<< c_dlu (fi nn) pa (fi (max 1 nn)) pperm)
    perm <- map fromIntegral <$> peekArray nn pperm
    pure (MkMatrix f, M.fromList M.Seq perm)
{-# NOINLINE lu #-}

-- | lu at e = Float (the reference's lu is generic over Fractional e,
-- Solve/LU.hs:121): the FP32 factorisation, with the same pivot rule.
luFloat :: forall n. KnownNat n
        => Matrix n n M.P Float -> (Matrix n n M.P Float, M.Array M.P M.Ix1 Int)
luFloat (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  allocaArray (max 1 nn) $ \pperm -> do
    ma <- M.thawS (M.computeAs M.S (M.delay a) :: M.Array M.S M.Ix2 Float)
    M.withPtr ma $ \pa -> () <$ (check =<< c_slu (fi nn) pa (fi (max 1 nn)) pperm)
    f <- M.computeAs M.P . M.delay <$> M.freezeS ma
    perm <- map fromIntegral <$> peekArray nn pperm
    pure (MkMatrix f, M.fromList M.Seq perm)
{-# NOINLINE luFloat #-}

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
mixed call = mixedAt @n call F32

-- | The factor precision of the refined solvers and resident factors:
-- FP64, FP32, or BF16 tiles with FP32 storage (LMX_PREC_*).
data Precision = F64 | F32 | BF16 deriving (Eq, Show)

precCode :: Precision -> CInt
precCode F64 = 0
precCode F32 = 1
precCode BF16 = 2

-- | luSolveMixed at a chosen factor precision (F32 or BF16).
luSolveMixedWith :: forall n. KnownNat n
                 => Precision -> Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
luSolveMixedWith = mixedAt @n c_dgesv_mixed

choleskySolveMixedWith :: forall n. KnownNat n
                       => Precision -> Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
choleskySolveMixedWith = mixedAt @n c_dposv_mixed

mixedAt :: forall n. KnownNat n
        => (CInt -> Int64 -> Int64 -> Ptr Double -> Int64 -> Ptr Double -> Int64
            -> Ptr Double -> Int64 -> Ptr () -> IO CInt)
        -> Precision -> Matrix n n M.P Double -> Vector n M.P Double -> Vector n M.P Double
mixedAt call prec (MkMatrix a) (MkVector b) = unsafePerformIO $ do
  let nn = dimVal @n
  ma <- M.thawS (toS2 a)
  mb <- M.thawS (toS1 b)
  mx <- M.thawS (toS1 b)
  allocaBytes 32 $ \info ->
    M.withPtr ma $ \pa -> M.withPtr mb $ \pb -> M.withPtr mx $ \px -> do
      rc <- call (precCode prec) (fi nn) 1 pa (fi (max 1 nn)) pb 1 px 1 (castPtr info)
      when (rc < 0) (() <$ check rc)
  MkVector . fromS1 <$> M.freezeS mx
{-# NOINLINE mixedAt #-}

-- ------------------------------------------------------------------ det, inverse iteration

-- | det (Solve/LU.hs:534-558): lu's factors, the product of U's diagonal in
-- order and the sign of the permutation.
det :: forall n. KnownNat n => Matrix n n M.P Double -> Double
det (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  ms <- M.thawS (toS2 a)
  alloca $ \pd -> do
    M.withPtr ms $ \pa -> () <$ (check =<< c_ddet (fi nn) pa (fi (max 1 nn)) pd)
    realToFrac <$> (peek pd :: IO Double)
{-# NOINLINE det #-}

-- | inverseIteration (Eigen/Power.hs:113-135): A - mu I is factored once on
-- the device; each step is a solve, a norm, a GEMV and a dot product there.
inverseIteration :: forall n. KnownNat n
                 => Matrix n n M.P Double -> Double -> Vector n M.P Double -> Int -> Double
                 -> (Double, Vector n M.P Double)
inverseIteration (MkMatrix a) mu (MkVector q0) maxIter tol = unsafePerformIO $ do
  let nn = dimVal @n
  ms <- M.thawS (toS2 a)
  mq0 <- M.thawS (toS1 q0)
  mq <- M.thawS (toS1 q0)
  lam <- alloca $ \pl -> alloca $ \pit ->
    M.withPtr ms $ \pa -> M.withPtr mq0 $ \pq0 -> M.withPtr mq $ \pq -> do
      _ <- check =<< c_dinverse_iteration (fi nn) pa (fi (max 1 nn)) (realToFrac mu) pq0
                                          (fromIntegral maxIter) (realToFrac tol) pl pq pit
      peek pl
  q <- fromS1 <$> M.freezeS mq
  pure (lam, MkVector q)
{-# NOINLINE inverseIteration #-}

-- ------------------------------------------------------------------ level 2, norms, SYRK

-- | gemv alpha A x beta y (BLAS/Level2.hs:92-100).  beta = 0 does not read y.
gemv :: forall m n. (KnownNat m, KnownNat n)
     => Double -> Matrix m n M.P Double -> Vector n M.P Double -> Double -> Vector m M.P Double
     -> Vector m M.P Double
gemv alpha (MkMatrix a) (MkVector x) beta (MkVector y) = unsafePerformIO $ do
  let mm = dimVal @m; nn = dimVal @n
  ms <- M.thawS (toS2 a)
  mx <- M.thawS (toS1 x)
  my <- M.thawS (toS1 y)
  M.withPtr ms $ \pa -> M.withPtr mx $ \px -> M.withPtr my $ \py ->
    () <$ (check =<< c_dgemv (fi mm) (fi nn) (realToFrac alpha) pa (fi (max 1 nn)) px (realToFrac beta) py)
  MkVector . fromS1 <$> M.freezeS my
{-# NOINLINE gemv #-}

-- | matvecP (BLAS/Level2.hs:132-140) and matvec (:123-130) at P Double.
matvecP :: forall m n. (KnownNat m, KnownNat n)
        => Matrix m n M.P Double -> Vector n M.P Double -> Vector m M.P Double
matvecP a x = gemv @m @n 1 a x 0 (MkVector (M.replicate M.Seq (M.Sz1 (dimVal @m)) 0))

matvec :: forall m n. (KnownNat m, KnownNat n)
       => Matrix m n M.P Double -> Vector n M.P Double -> Vector m M.P Double
matvec = matvecP @m @n

normWith :: forall m n. (KnownNat m, KnownNat n) => Char -> Matrix m n M.P Double -> Double
normWith which (MkMatrix a) = unsafePerformIO $ do
  let mm = dimVal @m; nn = dimVal @n
  ms <- M.thawS (toS2 a)
  alloca $ \pd -> do
    M.withPtr ms $ \pa -> () <$ (check =<< c_dnorm (chr which) (fi mm) (fi nn) pa (fi (max 1 nn)) pd)
    peek pd
{-# NOINLINE normWith #-}

-- | Norms.hs:124-130, :142-148, :110-112.
norm1, normInf, normFrob :: forall m n. (KnownNat m, KnownNat n) => Matrix m n M.P Double -> Double
norm1 = normWith @m @n '1'
normInf = normWith @m @n 'I'
normFrob = normWith @m @n 'F'

asColumn :: forall n. KnownNat n => Vector n M.P Double -> Matrix n 1 M.P Double
asColumn (MkVector x) = MkMatrix (M.resize' (M.Sz2 (dimVal @n) 1) x)

-- | Norms.hs:72-74, :83-85, :94-96.
vnorm1, vnorm2, vnormInf :: forall n. KnownNat n => Vector n M.P Double -> Double
vnorm1 = normWith @n @1 '1' . asColumn
vnorm2 = normWith @n @1 'F' . asColumn
vnormInf = normWith @n @1 'I' . asColumn

-- | matMulAtAP (BLAS/Level3.hs:353-373): A^T A, the whole symmetric matrix.
matMulAtAP :: forall m n. (KnownNat m, KnownNat n) => Matrix m n M.P Double -> Matrix n n M.P Double
matMulAtAP (MkMatrix a) = unsafePerformIO $ do
  let mm = dimVal @m; nn = dimVal @n
      c0 = M.replicate M.Seq (M.Sz2 nn nn) 0 :: M.Array M.P M.Ix2 Double
  ms <- M.thawS (toS2 a)
  c <- inPlace2 c0 $ \pc -> M.withPtr ms $ \pa ->
    () <$ (check =<< c_dsyrk (chr 'T') (fi nn) (fi mm) 1 pa (fi (max 1 nn)) 0 pc (fi (max 1 nn)))
  pure (MkMatrix c)
{-# NOINLINE matMulAtAP #-}

-- | A B on the BF16 matrix cores with FP32 accumulation; the operands hold
-- BF16 bit patterns (the upper half of an IEEE single).
matMulBF16 :: forall m k n. (KnownNat m, KnownNat k, KnownNat n)
           => Matrix m k M.P Word16 -> Matrix k n M.P Word16 -> Matrix m n M.P Float
matMulBF16 (MkMatrix a) (MkMatrix b) = unsafePerformIO $ do
  let mm = dimVal @m; kk = dimVal @k; nn = dimVal @n
  ma <- M.thawS (M.computeAs M.S (M.delay a) :: M.Array M.S M.Ix2 Word16)
  mb <- M.thawS (M.computeAs M.S (M.delay b) :: M.Array M.S M.Ix2 Word16)
  mc <- M.thawS (M.replicate M.Seq (M.Sz2 mm nn) 0 :: M.Array M.S M.Ix2 Float)
  M.withPtr ma $ \pa -> M.withPtr mb $ \pb -> M.withPtr mc $ \pc ->
    () <$ (check =<< c_gemm_bf16 (chr 'N') (chr 'N') (fi mm) (fi nn) (fi kk) 1 pa (fi (max 1 kk)) pb
                                (fi (max 1 nn)) 0 pc (fi (max 1 nn)))
  MkMatrix . M.computeAs M.P . M.delay <$> M.freezeS mc
{-# NOINLINE matMulBF16 #-}

-- ------------------------------------------------------------------ resident factors

-- | A factor of an n x n matrix kept in HBM (lmx_factor_*); released by the
-- garbage collector.  FP64 factors solve by substitution through the kept
-- triangular inverses; F32 / BF16 factors refine every solve to FP64
-- accuracy.
newtype Factor (n :: Nat) = Factor (ForeignPtr ())

mkFactor :: forall n. KnownNat n => CInt -> Precision -> Matrix n n M.P Double -> Factor n
mkFactor kind prec (MkMatrix a) = unsafePerformIO $ do
  let nn = dimVal @n
  ms <- M.thawS (toS2 a)
  h <- alloca $ \ph -> do
    M.withPtr ms $ \pa -> () <$ (check =<< c_factor_create ph kind (precCode prec) (fi nn) pa (fi (max 1 nn)))
    peek ph
  Factor <$> newForeignPtr h (void (c_factor_release h))
{-# NOINLINE mkFactor #-}

-- | An LU factor (partial pivoting) kept on the device.
luFactor :: forall n. KnownNat n => Precision -> Matrix n n M.P Double -> Factor n
luFactor = mkFactor @n 0

-- | A Cholesky factor kept on the device.
choleskyFactor :: forall n. KnownNat n => Precision -> Matrix n n M.P Double -> Factor n
choleskyFactor = mkFactor @n 1

-- | Many right-hand sides against one resident factor.
luSolveManyWith :: forall n k. (KnownNat n, KnownNat k)
                => Factor n -> Matrix n k M.P Double -> Matrix n k M.P Double
luSolveManyWith (Factor fp) (MkMatrix b) = unsafePerformIO $ do
  let nn = dimVal @n; kk = dimVal @k
  mb <- M.thawS (toS2 b)
  mx <- M.thawS (toS2 b)
  withForeignPtr fp $ \h -> allocaBytes 32 $ \info ->
    M.withPtr mb $ \pb -> M.withPtr mx $ \px -> do
      rc <- c_factor_solve h (fi kk) pb (fi (max 1 kk)) px (fi (max 1 kk)) (castPtr info)
      when (rc < 0) (() <$ check rc)
  MkMatrix . fromS2 <$> M.freezeS mx
{-# NOINLINE luSolveManyWith #-}

luSolveWith :: forall n. KnownNat n => Factor n -> Vector n M.P Double -> Vector n M.P Double
luSolveWith f (MkVector b) =
  let MkMatrix x = luSolveManyWith @n @1 f (MkMatrix (M.resize' (M.Sz2 (dimVal @n) 1) b))
  in MkVector (M.resize' (M.Sz1 (dimVal @n)) x)

choleskySolveWith :: forall n. KnownNat n => Factor n -> Vector n M.P Double -> Vector n M.P Double
choleskySolveWith = luSolveWith @n

fi :: Int -> Int64
fi = fromIntegral
